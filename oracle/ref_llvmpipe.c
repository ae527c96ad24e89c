/*
 * ref_llvmpipe.c -- runs the reference's own GLSL compute shaders on Mesa
 * llvmpipe through a headless DRI context. TEST INFRASTRUCTURE ONLY (golden
 * vectors and the reference CPU baseline); nothing in the product links it.
 *
 * The shader text is read at run time from the reference tree
 * (assets/shaders/voxelDraw.comp, voxelLighting.comp, voxelShared.comp) and
 * spliced the way src/DoonEngine/utility/shader.c:185-231 does: the shared
 * file is inserted right after the #version line. Optionally a harness-local
 * variant is compiled (never written anywhere) that
 *   - counts chunk-grid and voxel DDA trips and records the hit voxel index;
 *   - writes lighting results to a separate buffer and drops the in-shader
 *     sample-count increment and visible-bit clear, which race with other
 *     work groups (voxelLighting.comp:276-285); the caller applies them;
 *   - records the pre-quantisation diffuse/specular values.
 *
 * The GL driver is loaded from the system (swrast_dri.so); its entry points
 * come from libglapi's dispatch. Build: see oracle/build.py (output only into
 * oracle/_ref/).
 */
#define _GNU_SOURCE
#include <GL/internal/dri_interface.h>
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned int GLenum;
typedef unsigned int GLuint;
typedef int GLint;
typedef int GLsizei;
typedef float GLfloat;
typedef unsigned int GLbitfield;
typedef long GLsizeiptr;
typedef long GLintptr;
typedef unsigned char GLboolean;
typedef char GLchar;

#define GL_COMPUTE_SHADER 0x91B9
#define GL_COMPILE_STATUS 0x8B81
#define GL_LINK_STATUS 0x8B82
#define GL_SHADER_STORAGE_BUFFER 0x90D2
#define GL_DYNAMIC_DRAW 0x88E8
#define GL_TEXTURE_2D 0x0DE1
#define GL_RGBA32F 0x8814
#define GL_RGBA 0x1908
#define GL_FLOAT 0x1406
#define GL_WRITE_ONLY 0x88B9
#define GL_ALL_BARRIER_BITS 0xFFFFFFFF

static struct {
    GLuint (*CreateShader)(GLenum);
    void (*ShaderSource)(GLuint, GLsizei, const GLchar* const*, const GLint*);
    void (*CompileShader)(GLuint);
    void (*GetShaderiv)(GLuint, GLenum, GLint*);
    void (*GetShaderInfoLog)(GLuint, GLsizei, GLsizei*, GLchar*);
    GLuint (*CreateProgram)(void);
    void (*AttachShader)(GLuint, GLuint);
    void (*LinkProgram)(GLuint);
    void (*GetProgramiv)(GLuint, GLenum, GLint*);
    void (*GetProgramInfoLog)(GLuint, GLsizei, GLsizei*, GLchar*);
    void (*UseProgram)(GLuint);
    GLint (*GetUniformLocation)(GLuint, const GLchar*);
    void (*Uniform1ui)(GLint, GLuint);
    void (*Uniform1i)(GLint, GLint);
    void (*Uniform1f)(GLint, GLfloat);
    void (*Uniform3f)(GLint, GLfloat, GLfloat, GLfloat);
    void (*Uniform3ui)(GLint, GLuint, GLuint, GLuint);
    void (*UniformMatrix4fv)(GLint, GLsizei, GLboolean, const GLfloat*);
    void (*GenBuffers)(GLsizei, GLuint*);
    void (*DeleteBuffers)(GLsizei, const GLuint*);
    void (*BindBuffer)(GLenum, GLuint);
    void (*BufferData)(GLenum, GLsizeiptr, const void*, GLenum);
    void (*BindBufferBase)(GLenum, GLuint, GLuint);
    void (*GetBufferSubData)(GLenum, GLintptr, GLsizeiptr, void*);
    void (*GenTextures)(GLsizei, GLuint*);
    void (*DeleteTextures)(GLsizei, const GLuint*);
    void (*BindTexture)(GLenum, GLuint);
    void (*TexStorage2D)(GLenum, GLsizei, GLenum, GLsizei, GLsizei);
    void (*BindImageTexture)(GLuint, GLuint, GLint, GLboolean, GLint, GLenum, GLenum);
    void (*GetTexImage)(GLenum, GLint, GLenum, GLenum, void*);
    void (*DispatchCompute)(GLuint, GLuint, GLuint);
    void (*MemoryBarrier)(GLbitfield);
    void (*Finish)(void);
    GLenum (*GetError)(void);
    const unsigned char* (*GetString)(GLenum);
} gl;

static char g_err[4096];
static GLuint g_draw = 0, g_light = 0;
static int g_instrumented = 0;

const char* ref_error(void) { return g_err; }

static void dri_stub_info(__DRIdrawable* d, int* x, int* y, int* w, int* h, void* p)
{
    (void)d; (void)p;
    *x = *y = 0; *w = *h = 1;
}
static void dri_stub_put(__DRIdrawable* d, int op, int x, int y, int w, int h, char* data, void* p)
{
    (void)d; (void)op; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static void dri_stub_get(__DRIdrawable* d, int x, int y, int w, int h, char* data, void* p)
{
    (void)d; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static const __DRIswrastLoaderExtension g_loader = {{__DRI_SWRAST_LOADER, 1}, dri_stub_info, dri_stub_put, dri_stub_get,
                                                    NULL, NULL, NULL, NULL, NULL};
static const __DRIextension* g_loaderExts[] = {&g_loader.base, NULL};

static char* read_file(const char* path)
{
    FILE* f = fopen(path, "rb");
    if(!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* b = malloc((size_t)n + 1);
    if(fread(b, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(b); return NULL; }
    b[n] = 0;
    fclose(f);
    return b;
}

/* replace the first occurrence of `from` by `to`; returns a new string or NULL */
static char* replace1(char* src, const char* from, const char* to)
{
    char* p = strstr(src, from);
    if(!p) { snprintf(g_err, sizeof g_err, "instrumentation anchor not found: %.60s", from); free(src); return NULL; }
    size_t a = (size_t)(p - src), lf = strlen(from), lt = strlen(to), ls = strlen(src);
    char* r = malloc(ls - lf + lt + 1);
    memcpy(r, src, a);
    memcpy(r + a, to, lt);
    memcpy(r + a + lt, p + lf, ls - a - lf + 1);
    free(src);
    return r;
}

/* shader.c:185-231: the included file goes right after the #version line */
static char* splice(char* base, const char* inc)
{
    char* v = strstr(base, "#version");
    if(!v) { snprintf(g_err, sizeof g_err, "no #version"); return NULL; }
    size_t i = (size_t)(v - base);
    while(base[i] != '\n') i++;
    size_t lb = strlen(base), li = strlen(inc);
    char* r = malloc(lb + li + 1);
    memcpy(r, base, i);
    memcpy(r + i, inc, li);
    memcpy(r + i + li, base + i + 1, lb - i);
    free(base);
    return r;
}

static GLuint compile(const char* src)
{
    GLuint s = gl.CreateShader(GL_COMPUTE_SHADER);
    gl.ShaderSource(s, 1, &src, NULL);
    gl.CompileShader(s);
    GLint ok = 0;
    gl.GetShaderiv(s, GL_COMPILE_STATUS, &ok);
    if(!ok) { gl.GetShaderInfoLog(s, sizeof g_err, NULL, g_err); return 0; }
    GLuint p = gl.CreateProgram();
    gl.AttachShader(p, s);
    gl.LinkProgram(p);
    gl.GetProgramiv(p, GL_LINK_STATUS, &ok);
    if(!ok) { gl.GetProgramInfoLog(p, sizeof g_err, NULL, g_err); return 0; }
    return p;
}

static char* instrument_shared(char* s)
{
    s = s ? replace1(s, "bool enableRefraction = false;",
                     "bool enableRefraction = false;\nuint dbgMap = 0u; uint dbgVox = 0u; uint dbgHit = 0xFFFFFFFFu; uint dbgIdx = 0u;") : s;
    s = s ? replace1(s, "while(in_chunk_bounds(pos))\n\t{", "while(in_chunk_bounds(pos))\n\t{ dbgVox++;") : s;
    s = s ? replace1(s, "while(in_map_bounds(pos))\n\t{", "while(in_map_bounds(pos))\n\t{ dbgMap++;") : s;
    s = s ? replace1(s, "CompressedVoxel compressed = voxels[get_voxel_index(mapIndex, pos)];",
                     "dbgIdx = get_voxel_index(mapIndex, pos); CompressedVoxel compressed = voxels[dbgIdx];") : s;
    s = s ? replace1(s, "+ EPSILON);\n\t\t\t\treturn true;", "+ EPSILON);\n\t\t\t\tdbgHit = dbgIdx; return true;") : s;
    return s;
}

/* Creates the context and compiles both programs. instrument != 0 compiles the
 * harness-local variants described at the top of this file. */
int ref_init(const char* shaderDir, int instrument)
{
    static int ctxReady = 0;
    if(!ctxReady)
    {
        void* glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
        void* drv = dlopen("/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so", RTLD_NOW | RTLD_GLOBAL);
        if(!glapi || !drv) { snprintf(g_err, sizeof g_err, "dlopen: %s", dlerror()); return -1; }
        const __DRIextension** (*get)(void) = (const __DRIextension** (*)(void))dlsym(drv, "__driDriverGetExtensions_swrast");
        if(!get) { snprintf(g_err, sizeof g_err, "no swrast entry"); return -1; }
        const __DRIextension** ex = get();
        const __DRIcoreExtension* core = NULL;
        const __DRIswrastExtension* sw = NULL;
        for(int i = 0; ex[i]; i++)
        {
            if(!strcmp(ex[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension*)ex[i];
            if(!strcmp(ex[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension*)ex[i];
        }
        if(!core || !sw || sw->base.version < 4) { snprintf(g_err, sizeof g_err, "DRI extensions missing"); return -1; }
        const __DRIconfig** configs = NULL;
        __DRIscreen* scr = sw->createNewScreen2(0, g_loaderExts, ex, &configs, NULL);
        uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
        unsigned err = 0;
        __DRIcontext* ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE, configs ? configs[0] : NULL, NULL, 2,
                                                     attribs, &err, NULL);
        if(!ctx || !core->bindContext(ctx, NULL, NULL)) { snprintf(g_err, sizeof g_err, "context: %u", err); return -1; }
        void* (*gpa)(const char*) = (void* (*)(const char*))dlsym(glapi, "_glapi_get_proc_address");
#define L(n) *(void**)&gl.n = gpa("gl" #n); if(!gl.n) { snprintf(g_err, sizeof g_err, "missing gl" #n); return -1; }
        L(CreateShader) L(ShaderSource) L(CompileShader) L(GetShaderiv) L(GetShaderInfoLog) L(CreateProgram)
        L(AttachShader) L(LinkProgram) L(GetProgramiv) L(GetProgramInfoLog) L(UseProgram) L(GetUniformLocation)
        L(Uniform1ui) L(Uniform1i) L(Uniform1f) L(Uniform3f) L(Uniform3ui) L(UniformMatrix4fv) L(GenBuffers)
        L(DeleteBuffers) L(BindBuffer) L(BufferData) L(BindBufferBase) L(GetBufferSubData) L(GenTextures)
        L(DeleteTextures) L(BindTexture) L(TexStorage2D) L(BindImageTexture) L(GetTexImage) L(DispatchCompute)
        L(MemoryBarrier) L(Finish) L(GetError) L(GetString)
#undef L
        ctxReady = 1;
    }
    char path[1024];
    snprintf(path, sizeof path, "%s/voxelShared.comp", shaderDir);
    char* shared = read_file(path);
    snprintf(path, sizeof path, "%s/voxelDraw.comp", shaderDir);
    char* draw = read_file(path);
    snprintf(path, sizeof path, "%s/voxelLighting.comp", shaderDir);
    char* light = read_file(path);
    if(!shared || !draw || !light) { snprintf(g_err, sizeof g_err, "cannot read shaders in %s", shaderDir); return -2; }
    if(instrument)
    {
        shared = instrument_shared(shared);
        if(!shared) return -3;
        draw = replace1(draw, "uniform sampler2D colorSample;",
                        "layout(std430, binding = 5) buffer dbgBuffer { uint dbgOut[]; };\nuniform sampler2D colorSample;");
        draw = draw ? replace1(draw, "imageStore(colorOutput, coords, vec4(finalColor, finalDepth));",
                               "imageStore(colorOutput, coords, vec4(finalColor, finalDepth));\n"
                               "\tuint dbgP = (uint(coords.y) * uint(imageSize(colorOutput).x) + uint(coords.x)) * 3u;\n"
                               "\tdbgOut[dbgP] = dbgHit; dbgOut[dbgP + 1u] = dbgMap; dbgOut[dbgP + 2u] = dbgVox;") : draw;
        light = light ? replace1(light, "uniform float time;",
                                 "layout(std430, binding = 5) buffer dbgBuffer { float dbgVal[]; };\n"
                                 "layout(std430, binding = 6) buffer outBuffer { uvec4 voxelsOut[]; };\nuniform float time;") : light;
        light = light ? replace1(light, "\t//clamp to avoid compression errors:",
                                 "\tdbgVal[voxelIndex * 6u] = diffuseLight.x; dbgVal[voxelIndex * 6u + 1u] = diffuseLight.y;"
                                 " dbgVal[voxelIndex * 6u + 2u] = diffuseLight.z; dbgVal[voxelIndex * 6u + 3u] = specLight.x;"
                                 " dbgVal[voxelIndex * 6u + 4u] = specLight.y; dbgVal[voxelIndex * 6u + 5u] = specLight.z;\n"
                                 "\t//clamp to avoid compression errors:") : light;
        light = light ? replace1(light, "voxels[voxelIndex].albedo       =", "voxelsOut[voxelIndex].y =") : light;
        light = light ? replace1(light, "voxels[voxelIndex].specLight    =", "voxelsOut[voxelIndex].z =") : light;
        light = light ? replace1(light, "voxels[voxelIndex].diffuseLight =", "voxelsOut[voxelIndex].w =") : light;
        light = light ? replace1(light, "map[mapIndex].flags &= ~4;", "") : light;
        light = light ? replace1(light, "chunks[mapIndex].numIndirectSamples++;", "") : light;
        if(!draw || !light) return -3;
    }
    draw = splice(draw, shared);
    light = splice(light, shared);
    free(shared);
    if(!draw || !light) return -4;
    g_instrumented = instrument;
    g_draw = compile(draw);
    if(!g_draw) return -5;
    g_light = compile(light);
    if(!g_light) return -6;
    free(draw);
    free(light);
    return 0;
}

static GLuint ssbo(GLuint binding, const void* data, size_t bytes)
{
    GLuint b;
    gl.GenBuffers(1, &b);
    gl.BindBuffer(GL_SHADER_STORAGE_BUFFER, b);
    gl.BufferData(GL_SHADER_STORAGE_BUFFER, (GLsizeiptr)(bytes ? bytes : 16), data, GL_DYNAMIC_DRAW);
    gl.BindBufferBase(GL_SHADER_STORAGE_BUFFER, binding, b);
    return b;
}

static void readback(GLuint b, void* out, size_t bytes)
{
    gl.BindBuffer(GL_SHADER_STORAGE_BUFFER, b);
    gl.GetBufferSubData(GL_SHADER_STORAGE_BUFFER, 0, (GLsizeiptr)bytes, out);
}

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

typedef struct RefScene {
    const uint32_t* map; size_t tiles;      /* tiles * 3 words */
    const uint32_t* chunks;                 /* tiles * 24 words */
    const uint32_t* voxels; size_t voxelCount;
    const void* materials;                  /* 256 * 32 bytes */
    uint32_t mapSize[3];
    float sunStrength[3], ambient[3], skyBot[3], skyTop[3];
} RefScene;

static void common_uniforms(GLuint p, const RefScene* s)
{
    gl.Uniform3ui(gl.GetUniformLocation(p, "mapSize"), s->mapSize[0], s->mapSize[1], s->mapSize[2]);
    gl.Uniform1ui(gl.GetUniformLocation(p, "useCubemap"), 0);
    gl.Uniform3f(gl.GetUniformLocation(p, "skyGradientBot"), s->skyBot[0], s->skyBot[1], s->skyBot[2]);
    gl.Uniform3f(gl.GetUniformLocation(p, "skyGradientTop"), s->skyTop[0], s->skyTop[1], s->skyTop[2]);
    gl.Uniform3f(gl.GetUniformLocation(p, "sunStrength"), s->sunStrength[0], s->sunStrength[1], s->sunStrength[2]);
    gl.Uniform3f(gl.GetUniformLocation(p, "ambientStrength"), s->ambient[0], s->ambient[1], s->ambient[2]);
}

/* voxel.c:812-881 with the three inverse matrices given. Writes RGBA32F into
 * rgba (w*h*4), the map flags after the pass into mapOut (tiles*3), and, for
 * the instrumented variant, (hit voxel index, map trips, voxel trips) per
 * pixel into dbg. Returns seconds spent in dispatch + finish. */
double ref_draw(const RefScene* s, const float* invView, const float* invCenteredView, const float* invProj,
                uint32_t w, uint32_t h, uint32_t viewMode, float* rgba, uint32_t* mapOut, uint32_t* dbg)
{
    gl.UseProgram(g_draw);
    GLuint bm = ssbo(0, s->map, s->tiles * 12), bc = ssbo(1, s->chunks, s->tiles * 96),
           bmat = ssbo(2, s->materials, 256 * 32), bv = ssbo(4, s->voxels, s->voxelCount * 16),
           bd = ssbo(5, NULL, (size_t)w * h * 12);
    GLuint tex;
    gl.GenTextures(1, &tex);
    gl.BindTexture(GL_TEXTURE_2D, tex);
    gl.TexStorage2D(GL_TEXTURE_2D, 1, GL_RGBA32F, (GLsizei)w, (GLsizei)h);
    gl.BindImageTexture(0, tex, 0, 0, 0, GL_WRITE_ONLY, GL_RGBA32F);
    gl.Uniform1ui(gl.GetUniformLocation(g_draw, "composeRasterized"), 0);
    gl.Uniform1i(gl.GetUniformLocation(g_draw, "colorSample"), 0);
    gl.Uniform1i(gl.GetUniformLocation(g_draw, "depthSample"), 1);
    gl.Uniform1i(gl.GetUniformLocation(g_draw, "skyCubemap"), 2);
    gl.Uniform1ui(gl.GetUniformLocation(g_draw, "viewMode"), viewMode);
    gl.UniformMatrix4fv(gl.GetUniformLocation(g_draw, "invViewMat"), 1, 0, invView);
    gl.UniformMatrix4fv(gl.GetUniformLocation(g_draw, "invCenteredViewMat"), 1, 0, invCenteredView);
    gl.UniformMatrix4fv(gl.GetUniformLocation(g_draw, "invProjectionMat"), 1, 0, invProj);
    common_uniforms(g_draw, s);
    gl.Finish();
    double t0 = now();
    gl.DispatchCompute(w / 16, h / 16, 1);
    gl.MemoryBarrier(GL_ALL_BARRIER_BITS);
    gl.Finish();
    double t = now() - t0;
    if(rgba) gl.GetTexImage(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, rgba);
    if(mapOut) readback(bm, mapOut, s->tiles * 12);
    if(dbg) readback(bd, dbg, (size_t)w * h * 12);
    GLuint bufs[] = {bm, bc, bmat, bv, bd};
    gl.DeleteBuffers(5, bufs);
    gl.DeleteTextures(1, &tex);
    return t;
}

typedef struct RefLight {
    float time;
    uint32_t numDiffuseSamples, maxDiffuseSamples, diffuseBounceLimit, specBounceLimit;
    float sunDir[3], shadowSoftness, camPos[3];
} RefLight;

/* voxel.c:883-952. voxelsOut (voxelCount*4 words) receives the pool after the
 * pass (instrumented: the lit words merged into the input pool); dbg
 * (voxelCount*6 floats) the pre-quantisation values. Returns seconds. */
double ref_lighting(const RefScene* s, const RefLight* lp, const uint32_t* requests, uint32_t numRequests,
                    uint32_t* voxelsOut, uint32_t* mapOut, uint32_t* chunksOut, float* dbg)
{
    gl.UseProgram(g_light);
    GLuint bm = ssbo(0, s->map, s->tiles * 12), bc = ssbo(1, s->chunks, s->tiles * 96),
           bmat = ssbo(2, s->materials, 256 * 32), br = ssbo(3, requests, (size_t)numRequests * 4),
           bv = ssbo(4, s->voxels, s->voxelCount * 16), bd = ssbo(5, NULL, s->voxelCount * 24),
           bo = ssbo(6, s->voxels, s->voxelCount * 16);
    common_uniforms(g_light, s);
    gl.Uniform1i(gl.GetUniformLocation(g_light, "skyCubemap"), 2);
    gl.Uniform3f(gl.GetUniformLocation(g_light, "camPos"), lp->camPos[0], lp->camPos[1], lp->camPos[2]);
    gl.Uniform1f(gl.GetUniformLocation(g_light, "time"), lp->time);
    gl.Uniform1ui(gl.GetUniformLocation(g_light, "numDiffuseSamples"), lp->numDiffuseSamples);
    gl.Uniform1ui(gl.GetUniformLocation(g_light, "maxDiffuseSamples"), lp->maxDiffuseSamples);
    gl.Uniform1ui(gl.GetUniformLocation(g_light, "diffuseBounceLimit"), lp->diffuseBounceLimit);
    gl.Uniform1ui(gl.GetUniformLocation(g_light, "specularBounceLimit"), lp->specBounceLimit);
    gl.Uniform3f(gl.GetUniformLocation(g_light, "sunDir"), lp->sunDir[0], lp->sunDir[1], lp->sunDir[2]);
    gl.Uniform1f(gl.GetUniformLocation(g_light, "shadowSoftness"), lp->shadowSoftness);
    gl.Finish();
    double t0 = now();
    if(numRequests) gl.DispatchCompute(numRequests, 1, 1);
    gl.MemoryBarrier(GL_ALL_BARRIER_BITS);
    gl.Finish();
    double t = now() - t0;
    if(voxelsOut) readback(g_instrumented ? bo : bv, voxelsOut, s->voxelCount * 16);
    if(mapOut) readback(bm, mapOut, s->tiles * 12);
    if(chunksOut) readback(bc, chunksOut, s->tiles * 96);
    if(dbg) readback(bd, dbg, s->voxelCount * 24);
    GLuint bufs[] = {bm, bc, bmat, br, bv, bd, bo};
    gl.DeleteBuffers(7, bufs);
    return t;
}

const char* ref_renderer(void) { return (const char*)gl.GetString(0x1F01); }

/* Evaluates the shader hash fract(sin(s) * 43758.5453) * 2 - 1
 * (voxelLighting.comp:29-32) on llvmpipe for n seeds. */
int ref_rand(const float* seeds, float* out, uint32_t n)
{
    static GLuint prog = 0;
    if(!prog)
    {
        const char* src = "#version 430 core\n"
                          "layout(local_size_x = 64) in;\n"
                          "layout(std430, binding = 0) readonly buffer A { float s[]; };\n"
                          "layout(std430, binding = 1) writeonly buffer B { float r[]; };\n"
                          "uniform uint n;\n"
                          "float rand(float seed) { return fract(sin(seed) * 43758.5453) * 2.0 - 1.0; }\n"
                          "void main() { uint i = gl_GlobalInvocationID.x; if(i < n) r[i] = rand(s[i]); }\n";
        prog = compile(src);
        if(!prog) return -1;
    }
    gl.UseProgram(prog);
    GLuint a = ssbo(0, seeds, (size_t)n * 4), b = ssbo(1, NULL, (size_t)n * 4);
    gl.Uniform1ui(gl.GetUniformLocation(prog, "n"), n);
    gl.DispatchCompute((n + 63) / 64, 1, 1);
    gl.MemoryBarrier(GL_ALL_BARRIER_BITS);
    gl.Finish();
    readback(b, out, (size_t)n * 4);
    GLuint bufs[] = {a, b};
    gl.DeleteBuffers(2, bufs);
    return 0;
}
