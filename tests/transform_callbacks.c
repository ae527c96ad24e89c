/* Transform callbacks (VoxelTransformFunc, voxelShapes.h:28) for the sphere
 * edit tests: compiled by the tests and handed, as the same function pointer,
 * to the library's DN_shape_sphere and to the oracle's sphere edit. */
#include <doon/doon.h>

typedef struct CallbackLog {
    unsigned calls;
    float shapeMin[3], shapeMax[3];
    float inv[16];                  /* the last invTransform, m[col][row] */
} CallbackLog;

static unsigned char byte_of(float v) { return (unsigned char)((int)v & 255); }

/* A voxel that depends on every argument: its material and albedo on the
 * position and the incoming voxel, its normal on the position taken into the
 * shape's local space by invTransform (column-major m[col][row]). */
DNvoxel pattern(DNvec3 pos, DNvec3 normal, DNvoxel vox, DNvec3 shapeMin, DNvec3 shapeMax, DNmat4 inv, void* user)
{
    CallbackLog* log = (CallbackLog*)user;
    if(log)
    {
        log->calls++;
        log->shapeMin[0] = shapeMin.x; log->shapeMin[1] = shapeMin.y; log->shapeMin[2] = shapeMin.z;
        log->shapeMax[0] = shapeMax.x; log->shapeMax[1] = shapeMax.y; log->shapeMax[2] = shapeMax.z;
        for(int k = 0; k < 16; k++) log->inv[k] = inv.m[k / 4][k % 4];
    }
    float lx = inv.m[0][0] * pos.x + inv.m[1][0] * pos.y + inv.m[2][0] * pos.z + inv.m[3][0];
    float ly = inv.m[0][1] * pos.x + inv.m[1][1] * pos.y + inv.m[2][1] * pos.z + inv.m[3][1];
    float lz = inv.m[0][2] * pos.x + inv.m[1][2] * pos.y + inv.m[2][2] * pos.z + inv.m[3][2];
    DNvoxel out = vox;
    out.material = vox.material == DN_MATERIAL_EMPTY ? vox.material : (unsigned char)(1 + ((int)pos.x & 3));
    out.albedo.r = byte_of(vox.albedo.r + pos.x * 7.0f);
    out.albedo.g = byte_of(vox.albedo.g + pos.y * 11.0f);
    out.albedo.b = byte_of(vox.albedo.b + pos.z * 13.0f);
    out.normal.x = 0.5f * normal.x + 0.05f * lx;
    out.normal.y = 0.5f * normal.y + 0.05f * ly;
    out.normal.z = 0.5f * normal.z + 0.05f * lz;
    return out;
}

/* Keeps the incoming voxel and the computed normal: the same result as no callback. */
DNvoxel identity(DNvec3 pos, DNvec3 normal, DNvoxel vox, DNvec3 shapeMin, DNvec3 shapeMax, DNmat4 inv, void* user)
{
    (void)pos; (void)shapeMin; (void)shapeMax; (void)inv; (void)user;
    vox.normal = normal;
    return vox;
}
