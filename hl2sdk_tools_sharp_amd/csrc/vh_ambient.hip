// The leaf-ambient ray fan on gfx950: AmbientSampler.ComputeRayCube
// (SourceSharp.MapTools/Rad/Ambient/AmbientSampler.cs) for one light style.
//
// One block of 192 threads (three waves) per sample point; thread i < 162
// casts the ray along vertex normal i (Rad/Ambient/VertexNormals.cs) out to
// AmbientCube.RayLength, takes the closest hit through the BVH (the same
// exact triangle test and tie rule as vh_trace_closest) and turns it into the
// ray's light as RayAmbientLighting.Accumulate does
// (Rad/Ambient/RayAmbientLighting.cs): the hit face's average colour and the
// luxel under the hit, blended by the cone's radius at the hit
// (RemapValClamped 20..40), each tinted by the face's reflectivity, or the
// sky-ambient colour for a sky face.  The 162 colours meet in LDS, and six
// threads project them onto the box directions in ray order
// (AmbientCube.Project, Rad/Ambient/AmbientCube.cs).  Then the flagged world
// lights: a thread per light traces the segment to it, and the six threads
// add the unblocked ones in light order (AmbientCube.AddEmitSurfaceLights).
//
// The reference finds the hit by walking the BSP (AmbientRayTracer); here
// the hit is the closest-hit seam's, over the scene's triangles, whose caller
// ids carry the face index in their low 24 bits.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vh_internal.h"
#include "vh_traverse.h"

#pragma clang fp contract(off)

namespace {

using vh::kClosestInclusive;
using vh::kMaxDepth;
using vh::Ray;

constexpr int kFan = 162;       // VertexNormals.Count
constexpr int kFanBlock = 192;  // three waves; threads 162..191 only help project
constexpr float kRayLength = 32768.0f * 1.74f;  // AmbientCube.RayLength

// VertexNormals.All, in the reference's order (the anorms table).
__constant__ float kDirs[kFan * 3] = {
    -0.525731f, 0.000000f, 0.850651f, -0.442863f, 0.238856f, 0.864188f, -0.295242f, 0.000000f, 0.955423f,
    -0.309017f, 0.500000f, 0.809017f, -0.162460f, 0.262866f, 0.951056f, 0.000000f, 0.000000f, 1.000000f,
    0.000000f, 0.850651f, 0.525731f, -0.147621f, 0.716567f, 0.681718f, 0.147621f, 0.716567f, 0.681718f,
    0.000000f, 0.525731f, 0.850651f, 0.309017f, 0.500000f, 0.809017f, 0.525731f, 0.000000f, 0.850651f,
    0.295242f, 0.000000f, 0.955423f, 0.442863f, 0.238856f, 0.864188f, 0.162460f, 0.262866f, 0.951056f,
    -0.681718f, 0.147621f, 0.716567f, -0.809017f, 0.309017f, 0.500000f, -0.587785f, 0.425325f, 0.688191f,
    -0.850651f, 0.525731f, 0.000000f, -0.864188f, 0.442863f, 0.238856f, -0.716567f, 0.681718f, 0.147621f,
    -0.688191f, 0.587785f, 0.425325f, -0.500000f, 0.809017f, 0.309017f, -0.238856f, 0.864188f, 0.442863f,
    -0.425325f, 0.688191f, 0.587785f, -0.716567f, 0.681718f, -0.147621f, -0.500000f, 0.809017f, -0.309017f,
    -0.525731f, 0.850651f, 0.000000f, 0.000000f, 0.850651f, -0.525731f, -0.238856f, 0.864188f, -0.442863f,
    0.000000f, 0.955423f, -0.295242f, -0.262866f, 0.951056f, -0.162460f, 0.000000f, 1.000000f, 0.000000f,
    0.000000f, 0.955423f, 0.295242f, -0.262866f, 0.951056f, 0.162460f, 0.238856f, 0.864188f, 0.442863f,
    0.262866f, 0.951056f, 0.162460f, 0.500000f, 0.809017f, 0.309017f, 0.238856f, 0.864188f, -0.442863f,
    0.262866f, 0.951056f, -0.162460f, 0.500000f, 0.809017f, -0.309017f, 0.850651f, 0.525731f, 0.000000f,
    0.716567f, 0.681718f, 0.147621f, 0.716567f, 0.681718f, -0.147621f, 0.525731f, 0.850651f, 0.000000f,
    0.425325f, 0.688191f, 0.587785f, 0.864188f, 0.442863f, 0.238856f, 0.688191f, 0.587785f, 0.425325f,
    0.809017f, 0.309017f, 0.500000f, 0.681718f, 0.147621f, 0.716567f, 0.587785f, 0.425325f, 0.688191f,
    0.955423f, 0.295242f, 0.000000f, 1.000000f, 0.000000f, 0.000000f, 0.951056f, 0.162460f, 0.262866f,
    0.850651f, -0.525731f, 0.000000f, 0.955423f, -0.295242f, 0.000000f, 0.864188f, -0.442863f, 0.238856f,
    0.951056f, -0.162460f, 0.262866f, 0.809017f, -0.309017f, 0.500000f, 0.681718f, -0.147621f, 0.716567f,
    0.850651f, 0.000000f, 0.525731f, 0.864188f, 0.442863f, -0.238856f, 0.809017f, 0.309017f, -0.500000f,
    0.951056f, 0.162460f, -0.262866f, 0.525731f, 0.000000f, -0.850651f, 0.681718f, 0.147621f, -0.716567f,
    0.681718f, -0.147621f, -0.716567f, 0.850651f, 0.000000f, -0.525731f, 0.809017f, -0.309017f, -0.500000f,
    0.864188f, -0.442863f, -0.238856f, 0.951056f, -0.162460f, -0.262866f, 0.147621f, 0.716567f, -0.681718f,
    0.309017f, 0.500000f, -0.809017f, 0.425325f, 0.688191f, -0.587785f, 0.442863f, 0.238856f, -0.864188f,
    0.587785f, 0.425325f, -0.688191f, 0.688191f, 0.587785f, -0.425325f, -0.147621f, 0.716567f, -0.681718f,
    -0.309017f, 0.500000f, -0.809017f, 0.000000f, 0.525731f, -0.850651f, -0.525731f, 0.000000f, -0.850651f,
    -0.442863f, 0.238856f, -0.864188f, -0.295242f, 0.000000f, -0.955423f, -0.162460f, 0.262866f, -0.951056f,
    0.000000f, 0.000000f, -1.000000f, 0.295242f, 0.000000f, -0.955423f, 0.162460f, 0.262866f, -0.951056f,
    -0.442863f, -0.238856f, -0.864188f, -0.309017f, -0.500000f, -0.809017f, -0.162460f, -0.262866f, -0.951056f,
    0.000000f, -0.850651f, -0.525731f, -0.147621f, -0.716567f, -0.681718f, 0.147621f, -0.716567f, -0.681718f,
    0.000000f, -0.525731f, -0.850651f, 0.309017f, -0.500000f, -0.809017f, 0.442863f, -0.238856f, -0.864188f,
    0.162460f, -0.262866f, -0.951056f, 0.238856f, -0.864188f, -0.442863f, 0.500000f, -0.809017f, -0.309017f,
    0.425325f, -0.688191f, -0.587785f, 0.716567f, -0.681718f, -0.147621f, 0.688191f, -0.587785f, -0.425325f,
    0.587785f, -0.425325f, -0.688191f, 0.000000f, -0.955423f, -0.295242f, 0.000000f, -1.000000f, 0.000000f,
    0.262866f, -0.951056f, -0.162460f, 0.000000f, -0.850651f, 0.525731f, 0.000000f, -0.955423f, 0.295242f,
    0.238856f, -0.864188f, 0.442863f, 0.262866f, -0.951056f, 0.162460f, 0.500000f, -0.809017f, 0.309017f,
    0.716567f, -0.681718f, 0.147621f, 0.525731f, -0.850651f, 0.000000f, -0.238856f, -0.864188f, -0.442863f,
    -0.500000f, -0.809017f, -0.309017f, -0.262866f, -0.951056f, -0.162460f, -0.850651f, -0.525731f, 0.000000f,
    -0.716567f, -0.681718f, -0.147621f, -0.716567f, -0.681718f, 0.147621f, -0.525731f, -0.850651f, 0.000000f,
    -0.500000f, -0.809017f, 0.309017f, -0.238856f, -0.864188f, 0.442863f, -0.262866f, -0.951056f, 0.162460f,
    -0.864188f, -0.442863f, 0.238856f, -0.809017f, -0.309017f, 0.500000f, -0.688191f, -0.587785f, 0.425325f,
    -0.681718f, -0.147621f, 0.716567f, -0.442863f, -0.238856f, 0.864188f, -0.587785f, -0.425325f, 0.688191f,
    -0.309017f, -0.500000f, 0.809017f, -0.147621f, -0.716567f, 0.681718f, -0.425325f, -0.688191f, 0.587785f,
    -0.162460f, -0.262866f, 0.951056f, 0.442863f, -0.238856f, 0.864188f, 0.162460f, -0.262866f, 0.951056f,
    0.309017f, -0.500000f, 0.809017f, 0.147621f, -0.716567f, 0.681718f, 0.000000f, -0.525731f, 0.850651f,
    0.425325f, -0.688191f, 0.587785f, 0.587785f, -0.425325f, 0.688191f, 0.688191f, -0.587785f, 0.425325f,
    -0.955423f, 0.295242f, 0.000000f, -0.951056f, 0.162460f, 0.262866f, -1.000000f, 0.000000f, 0.000000f,
    -0.850651f, 0.000000f, 0.525731f, -0.955423f, -0.295242f, 0.000000f, -0.951056f, -0.162460f, 0.262866f,
    -0.864188f, 0.442863f, -0.238856f, -0.951056f, 0.162460f, -0.262866f, -0.809017f, 0.309017f, -0.500000f,
    -0.864188f, -0.442863f, -0.238856f, -0.951056f, -0.162460f, -0.262866f, -0.809017f, -0.309017f, -0.500000f,
    -0.681718f, 0.147621f, -0.716567f, -0.681718f, -0.147621f, -0.716567f, -0.850651f, 0.000000f, -0.525731f,
    -0.688191f, 0.587785f, -0.425325f, -0.587785f, 0.425325f, -0.688191f, -0.425325f, 0.688191f, -0.587785f,
    -0.425325f, -0.688191f, -0.587785f, -0.587785f, -0.425325f, -0.688191f, -0.688191f, -0.587785f, -0.425325f,
};

// The same table for the host (vh_ambient_directions).
const float kHostDirs[kFan * 3] = {
    -0.525731f, 0.000000f, 0.850651f, -0.442863f, 0.238856f, 0.864188f, -0.295242f, 0.000000f, 0.955423f,
    -0.309017f, 0.500000f, 0.809017f, -0.162460f, 0.262866f, 0.951056f, 0.000000f, 0.000000f, 1.000000f,
    0.000000f, 0.850651f, 0.525731f, -0.147621f, 0.716567f, 0.681718f, 0.147621f, 0.716567f, 0.681718f,
    0.000000f, 0.525731f, 0.850651f, 0.309017f, 0.500000f, 0.809017f, 0.525731f, 0.000000f, 0.850651f,
    0.295242f, 0.000000f, 0.955423f, 0.442863f, 0.238856f, 0.864188f, 0.162460f, 0.262866f, 0.951056f,
    -0.681718f, 0.147621f, 0.716567f, -0.809017f, 0.309017f, 0.500000f, -0.587785f, 0.425325f, 0.688191f,
    -0.850651f, 0.525731f, 0.000000f, -0.864188f, 0.442863f, 0.238856f, -0.716567f, 0.681718f, 0.147621f,
    -0.688191f, 0.587785f, 0.425325f, -0.500000f, 0.809017f, 0.309017f, -0.238856f, 0.864188f, 0.442863f,
    -0.425325f, 0.688191f, 0.587785f, -0.716567f, 0.681718f, -0.147621f, -0.500000f, 0.809017f, -0.309017f,
    -0.525731f, 0.850651f, 0.000000f, 0.000000f, 0.850651f, -0.525731f, -0.238856f, 0.864188f, -0.442863f,
    0.000000f, 0.955423f, -0.295242f, -0.262866f, 0.951056f, -0.162460f, 0.000000f, 1.000000f, 0.000000f,
    0.000000f, 0.955423f, 0.295242f, -0.262866f, 0.951056f, 0.162460f, 0.238856f, 0.864188f, 0.442863f,
    0.262866f, 0.951056f, 0.162460f, 0.500000f, 0.809017f, 0.309017f, 0.238856f, 0.864188f, -0.442863f,
    0.262866f, 0.951056f, -0.162460f, 0.500000f, 0.809017f, -0.309017f, 0.850651f, 0.525731f, 0.000000f,
    0.716567f, 0.681718f, 0.147621f, 0.716567f, 0.681718f, -0.147621f, 0.525731f, 0.850651f, 0.000000f,
    0.425325f, 0.688191f, 0.587785f, 0.864188f, 0.442863f, 0.238856f, 0.688191f, 0.587785f, 0.425325f,
    0.809017f, 0.309017f, 0.500000f, 0.681718f, 0.147621f, 0.716567f, 0.587785f, 0.425325f, 0.688191f,
    0.955423f, 0.295242f, 0.000000f, 1.000000f, 0.000000f, 0.000000f, 0.951056f, 0.162460f, 0.262866f,
    0.850651f, -0.525731f, 0.000000f, 0.955423f, -0.295242f, 0.000000f, 0.864188f, -0.442863f, 0.238856f,
    0.951056f, -0.162460f, 0.262866f, 0.809017f, -0.309017f, 0.500000f, 0.681718f, -0.147621f, 0.716567f,
    0.850651f, 0.000000f, 0.525731f, 0.864188f, 0.442863f, -0.238856f, 0.809017f, 0.309017f, -0.500000f,
    0.951056f, 0.162460f, -0.262866f, 0.525731f, 0.000000f, -0.850651f, 0.681718f, 0.147621f, -0.716567f,
    0.681718f, -0.147621f, -0.716567f, 0.850651f, 0.000000f, -0.525731f, 0.809017f, -0.309017f, -0.500000f,
    0.864188f, -0.442863f, -0.238856f, 0.951056f, -0.162460f, -0.262866f, 0.147621f, 0.716567f, -0.681718f,
    0.309017f, 0.500000f, -0.809017f, 0.425325f, 0.688191f, -0.587785f, 0.442863f, 0.238856f, -0.864188f,
    0.587785f, 0.425325f, -0.688191f, 0.688191f, 0.587785f, -0.425325f, -0.147621f, 0.716567f, -0.681718f,
    -0.309017f, 0.500000f, -0.809017f, 0.000000f, 0.525731f, -0.850651f, -0.525731f, 0.000000f, -0.850651f,
    -0.442863f, 0.238856f, -0.864188f, -0.295242f, 0.000000f, -0.955423f, -0.162460f, 0.262866f, -0.951056f,
    0.000000f, 0.000000f, -1.000000f, 0.295242f, 0.000000f, -0.955423f, 0.162460f, 0.262866f, -0.951056f,
    -0.442863f, -0.238856f, -0.864188f, -0.309017f, -0.500000f, -0.809017f, -0.162460f, -0.262866f, -0.951056f,
    0.000000f, -0.850651f, -0.525731f, -0.147621f, -0.716567f, -0.681718f, 0.147621f, -0.716567f, -0.681718f,
    0.000000f, -0.525731f, -0.850651f, 0.309017f, -0.500000f, -0.809017f, 0.442863f, -0.238856f, -0.864188f,
    0.162460f, -0.262866f, -0.951056f, 0.238856f, -0.864188f, -0.442863f, 0.500000f, -0.809017f, -0.309017f,
    0.425325f, -0.688191f, -0.587785f, 0.716567f, -0.681718f, -0.147621f, 0.688191f, -0.587785f, -0.425325f,
    0.587785f, -0.425325f, -0.688191f, 0.000000f, -0.955423f, -0.295242f, 0.000000f, -1.000000f, 0.000000f,
    0.262866f, -0.951056f, -0.162460f, 0.000000f, -0.850651f, 0.525731f, 0.000000f, -0.955423f, 0.295242f,
    0.238856f, -0.864188f, 0.442863f, 0.262866f, -0.951056f, 0.162460f, 0.500000f, -0.809017f, 0.309017f,
    0.716567f, -0.681718f, 0.147621f, 0.525731f, -0.850651f, 0.000000f, -0.238856f, -0.864188f, -0.442863f,
    -0.500000f, -0.809017f, -0.309017f, -0.262866f, -0.951056f, -0.162460f, -0.850651f, -0.525731f, 0.000000f,
    -0.716567f, -0.681718f, -0.147621f, -0.716567f, -0.681718f, 0.147621f, -0.525731f, -0.850651f, 0.000000f,
    -0.500000f, -0.809017f, 0.309017f, -0.238856f, -0.864188f, 0.442863f, -0.262866f, -0.951056f, 0.162460f,
    -0.864188f, -0.442863f, 0.238856f, -0.809017f, -0.309017f, 0.500000f, -0.688191f, -0.587785f, 0.425325f,
    -0.681718f, -0.147621f, 0.716567f, -0.442863f, -0.238856f, 0.864188f, -0.587785f, -0.425325f, 0.688191f,
    -0.309017f, -0.500000f, 0.809017f, -0.147621f, -0.716567f, 0.681718f, -0.425325f, -0.688191f, 0.587785f,
    -0.162460f, -0.262866f, 0.951056f, 0.442863f, -0.238856f, 0.864188f, 0.162460f, -0.262866f, 0.951056f,
    0.309017f, -0.500000f, 0.809017f, 0.147621f, -0.716567f, 0.681718f, 0.000000f, -0.525731f, 0.850651f,
    0.425325f, -0.688191f, 0.587785f, 0.587785f, -0.425325f, 0.688191f, 0.688191f, -0.587785f, 0.425325f,
    -0.955423f, 0.295242f, 0.000000f, -0.951056f, 0.162460f, 0.262866f, -1.000000f, 0.000000f, 0.000000f,
    -0.850651f, 0.000000f, 0.525731f, -0.955423f, -0.295242f, 0.000000f, -0.951056f, -0.162460f, 0.262866f,
    -0.864188f, 0.442863f, -0.238856f, -0.951056f, 0.162460f, -0.262866f, -0.809017f, 0.309017f, -0.500000f,
    -0.864188f, -0.442863f, -0.238856f, -0.951056f, -0.162460f, -0.262866f, -0.809017f, -0.309017f, -0.500000f,
    -0.681718f, 0.147621f, -0.716567f, -0.681718f, -0.147621f, -0.716567f, -0.850651f, 0.000000f, -0.525731f,
    -0.688191f, 0.587785f, -0.425325f, -0.587785f, 0.425325f, -0.688191f, -0.425325f, 0.688191f, -0.587785f,
    -0.425325f, -0.688191f, -0.587785f, -0.587785f, -0.425325f, -0.688191f, -0.688191f, -0.587785f, -0.425325f,
};

struct AmbientFaceDev {
    float s[4], t[4];
    float min_s, min_t;
    int32_t smax, tmax, light_ofs, sky;
    float refl[3];
};

// TexLightToLinear: mantissa * (float)(2^e / 255.0) (Rad/Ambient/StockLightColor.cs:69-78, 262-271).
__device__ __forceinline__ float3 decode(uint32_t w)
{
    int e = (int)(int8_t)(uint8_t)(w >> 24);
    float p = (float)(ldexp(1.0, e) / 255.0);
    return make_float3((float)(w & 255u) * p, (float)((w >> 8) & 255u) * p, (float)((w >> 16) & 255u) * p);
}

// A world light flagged into the cubes.
struct CubeLight {
    float o[3], n[3], i[3], radius;
};

struct AmbientArgs {
    const float4 *nodes, *tris;
    const CubeLight *lights;
    int32_t nlights;
    const AmbientFaceDev *faces;
    int32_t nfaces;
    const float3 *avg;        // AverageTinted per face
    const uint32_t *lump;
    int32_t has_sky;
    float sky0, sky1, sky2;
    float tan_theta;
    const float *points;
    int64_t npoints;
    float *cubes;
};

__global__ __launch_bounds__(kFanBlock) void k_ambient(AmbientArgs a)
{
    __shared__ int stack_mem[kMaxDepth * kFanBlock];
    __shared__ float rad[kFan * 3];
    __shared__ unsigned char vis[kFanBlock];
    const int i = threadIdx.x;
    const int64_t pidx = blockIdx.x;
    const float px = a.points[pidx * 3], py = a.points[pidx * 3 + 1], pz = a.points[pidx * 3 + 2];
    if (i < kFan) {
        Ray r;
        r.ox = px;
        r.oy = py;
        r.oz = pz;
        r.dx = kDirs[i * 3] * kRayLength;
        r.dy = kDirs[i * 3 + 1] * kRayLength;
        r.dz = kDirs[i * 3 + 2] * kRayLength;
        r.reach = 1.0f;
        float cr = 0.0f, cg = 0.0f, cb = 0.0f;
        int32_t face = -1;
        float frac = 1.0f;
        if (ray_finite(r)) {
            float t;
            int32_t id;
            int32_t h = trace_one<kClosestInclusive, false, kFanBlock>(a.nodes, a.tris, stack_mem + i, r,
                                                                        0.0f, -1, t, id);
            if (h >= 0) {
                face = id & 0x00FFFFFF;
                frac = t;
            }
        }
        if (face >= 0 && face < a.nfaces) {
            const AmbientFaceDev &F = a.faces[face];
            const bool sky = F.sky != 0;
            // the cone's radius where it meets the surface
            float len = sqrtf(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
            float dist = len * a.tan_theta * frac;
            // RemapValClamped(dist, 20, 40, 0, 1)
            float cv = (dist - 20.0f) / (40.0f - 20.0f);
            cv = fminf(fmaxf(cv, 0.0f), 1.0f);
            float s_avg = 0.0f + (1.0f - 0.0f) * cv;
            if (sky) s_avg = 1.0f;  // a sky hit has no luxel
            const float s_smp = 1.0f - s_avg;
            if (s_avg != 0.0f) {
                // AddFaceAverage
                if (sky) {
                    if (a.has_sky) {
                        cr = cr + a.sky0 * s_avg;
                        cg = cg + a.sky1 * s_avg;
                        cb = cb + a.sky2 * s_avg;
                    }
                } else {
                    float3 v = a.avg[face];
                    cr = cr + v.x * s_avg;
                    cg = cg + v.y * s_avg;
                    cb = cb + v.z * s_avg;
                }
            }
            if (s_smp != 0.0f && F.light_ofs != -1) {
                // AddFacePointSample: the luxel under the hit, truncated and clamped
                float hx = r.ox + frac * r.dx, hy = r.oy + frac * r.dy, hz = r.oz + frac * r.dz;
                float sc = hx * F.s[0] + hy * F.s[1] + hz * F.s[2] + F.s[3];
                float tc = hx * F.t[0] + hy * F.t[1] + hz * F.t[2] + F.t[3];
                float ds = sc - F.min_s, dt = tc - F.min_t;
                int32_t di = ds != ds ? 0 : (int32_t)ds;
                int32_t ti = dt != dt ? 0 : (int32_t)dt;
                di = min(max(di, 0), max(F.smax - 1, 0));
                ti = min(max(ti, 0), max(F.tmax - 1, 0));
                int64_t index = (int64_t)ti * F.smax + di;
                float3 c = decode(a.lump[(F.light_ofs >> 2) + index]);
                c.x = c.x * F.refl[0];
                c.y = c.y * F.refl[1];
                c.z = c.z * F.refl[2];
                cr = cr + c.x * s_smp;
                cg = cg + c.y * s_smp;
                cb = cb + c.z * s_smp;
            }
        }
        rad[i * 3] = cr;
        rad[i * 3 + 1] = cg;
        rad[i * 3 + 2] = cb;
    }
    __syncthreads();
    // AmbientCube.Project for box direction j = i, rays in order
    const int j = i;
    const float bx = j == 0 ? 1.0f : (j == 1 ? -1.0f : 0.0f);
    const float by = j == 2 ? 1.0f : (j == 3 ? -1.0f : 0.0f);
    const float bz = j == 4 ? 1.0f : (j == 5 ? -1.0f : 0.0f);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (j < 6) {
        float t = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        for (int k = 0; k < kFan; k++) {
            float c = kDirs[k * 3] * bx + kDirs[k * 3 + 1] * by + kDirs[k * 3 + 2] * bz;
            if (c > 0.0f) {
                t = t + c;
                sr = sr + rad[k * 3] * c;
                sg = sg + rad[k * 3 + 1] * c;
                sb = sb + rad[k * 3 + 2] * c;
            }
        }
        float inv = 1.0f / t;
        cr = sr * inv;
        cg = sg * inv;
        cb = sb * inv;
    }
    // AddSurfaceLights: each flagged light whose segment is open
    for (int32_t base = 0; base < a.nlights; base += kFanBlock) {
        __syncthreads();
        const int32_t l = base + i;
        if (l < a.nlights) {
            // Ray.Segment (Tracing/IRayTracer.cs:71-88), traced as TestLine
            const CubeLight &L = a.lights[l];
            float dx = L.o[0] - px, dy = L.o[1] - py, dz = L.o[2] - pz;
            float len = sqrtf(dx * dx + dy * dy + dz * dz);
            float inv = 1.0f / len;
            Ray r;
            r.ox = px;
            r.oy = py;
            r.oz = pz;
            r.dx = dx * inv;
            r.dy = dy * inv;
            r.dz = dz * inv;
            r.reach = len;
            bool open = true;
            if (vh::ray_finite(r)) {
                float t;
                int32_t id;
                open = vh::trace_one<vh::kAny, false, kFanBlock>(a.nodes, a.tris, stack_mem + i, r, 0.0f, -1,
                                                                  t, id) < 0;
            }
            vis[i] = open;
        }
        __syncthreads();
        if (j < 6) {
            const int32_t m = min(kFanBlock, a.nlights - base);
            for (int32_t k = 0; k < m; k++) {
                if (!vis[k]) continue;
                const CubeLight &L = a.lights[base + k];
                // AmbientCube.UnoccludedScale: DistanceFalloff times WorldLightAngle
                float dx = L.o[0] - px, dy = L.o[1] - py, dz = L.o[2] - pz;
                float r2 = dx * dx + dy * dy + dz * dz;
                float dscale;
                if (L.radius != 0.0f && r2 > L.radius * L.radius)
                    dscale = 0.0f;
                else
                    dscale = 1.0f / (r2 < 1.0f ? 1.0f : r2);
                float len = sqrtf(r2);
                float nx = 0.0f, ny = 0.0f, nz = 0.0f;
                if (len != 0.0f) {
                    nx = dx / len;
                    ny = dy / len;
                    nz = dz / len;
                }
                float dot = nx * nx + ny * ny + nz * nz;
                float angle = 0.0f;
                if (!(dot < 0.0f)) {
                    float dot2 = -(nx * L.n[0] + ny * L.n[1] + nz * L.n[2]);
                    if (!((double)dot2 <= 0.01)) angle = dot * dot2;
                }
                float ratio = dscale * angle * 1.0f;
                if (ratio == 0.0f) continue;
                float tj = bx * nx + by * ny + bz * nz;
                if (tj > 0.0f) {
                    float w = tj * ratio;
                    cr = cr + L.i[0] * w;
                    cg = cg + L.i[1] * w;
                    cb = cb + L.i[2] * w;
                }
            }
        }
    }
    if (j < 6) {
        float *o = a.cubes + (pidx * 6 + j) * 3;
        o[0] = cr;
        o[1] = cg;
        o[2] = cb;
    }
}

// AverageTinted per face (AmbientScene.BuildShades): the average word before
// the face's luxels decoded, times the reflectivity; zero for sky faces and
// faces without a lightmap.
__global__ void k_ambient_avg(const AmbientFaceDev *__restrict__ faces, int32_t n,
                              const uint32_t *__restrict__ lump, float3 *__restrict__ avg)
{
    int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const AmbientFaceDev &F = faces[f];
    float3 v = make_float3(0.0f, 0.0f, 0.0f);
    if (!F.sky && F.light_ofs >= 4) {
        float3 c = decode(lump[(F.light_ofs >> 2) - 1]);
        v = make_float3(c.x * F.refl[0], c.y * F.refl[1], c.z * F.refl[2]);
    }
    avg[f] = v;
}

}  // namespace

struct AmbientData {
    int32_t nfaces = 0;
    int64_t lump_bytes = 0;
    bool own_lump = false;
    int32_t has_sky = 0;
    float sky[3] = {0, 0, 0};
    int32_t nflag = 0;
    vh::DevBuf faces, avg, lump, points, cubes, lights;
};

namespace vh {
void ambient_release(vh_ctx *c)
{
    AmbientData *D = c->ambient;
    if (!D) return;
    for (DevBuf *b : {&D->faces, &D->avg, &D->lump, &D->points, &D->cubes, &D->lights}) b->release();
    delete D;
    c->ambient = nullptr;
}
}  // namespace vh

using namespace vh;

extern "C" {

int vh_ambient_directions(float *out)
{
    clear_error();
    if (!out) {
        set_error("NULL output");
        return VH_EINVAL;
    }
    std::memcpy(out, kHostDirs, sizeof kHostDirs);
    return VH_OK;
}

int vh_ambient_load(vh_ctx *c, const vh_ambient_desc *d)
{
    clear_error();
    if (!c || !d || d->nfaces < 0 || (d->nfaces > 0 && !d->faces) || d->lump_bytes < 0 ||
        (d->lump && d->lump_bytes % 4 != 0) || d->nlights < 0 || (d->nlights > 0 && !d->lights)) {
        set_error("bad leaf-ambient descriptor");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int64_t lump_bytes = d->lump_bytes;
    const void *dev_lump = nullptr;
    if (!d->lump) {
        FinalLump fl;
        if (!final_lump(c, &fl)) {
            set_error("no lump: pass one, or run vh_final_light first");
            return VH_ESTATE;
        }
        dev_lump = fl.lump;
        lump_bytes = fl.bytes;
    }
    std::vector<AmbientFaceDev> fd((size_t)std::max(d->nfaces, 1));
    for (int32_t f = 0; f < d->nfaces; f++) {
        const vh_ambient_face &x = d->faces[f];
        if (x.smax < 0 || x.tmax < 0 || (x.light_ofs != -1 && (x.light_ofs < 4 || x.light_ofs % 4 != 0 ||
            (int64_t)x.light_ofs + 4 * (int64_t)x.smax * x.tmax > lump_bytes ||
            (int64_t)x.smax * x.tmax == 0))) {
            set_error("face " + std::to_string(f) + ": its lightmap does not fit the lump");
            return VH_EINVAL;
        }
        AmbientFaceDev &o = fd[f];
        std::memcpy(o.s, x.s, sizeof o.s);
        std::memcpy(o.t, x.t, sizeof o.t);
        o.min_s = x.min_s;
        o.min_t = x.min_t;
        o.smax = x.smax;
        o.tmax = x.tmax;
        o.light_ofs = x.light_ofs;
        o.sky = x.sky != 0;
        std::memcpy(o.refl, x.reflectivity, sizeof o.refl);
    }
    // LeafAmbientSurfaceLights.Classify, Correct side: surface lights of
    // style 0 whose brightest channel times InvRSquared((0, 0, 512)) is under 0.005
    std::vector<CubeLight> flagged;
    const float ratio512 = 1.0f / (512.0f * 512.0f);
    for (int32_t l = 0; l < d->nlights; l++) {
        const vh_ambient_light &x = d->lights[l];
        if (x.type != VH_EMIT_SURFACE || x.style != 0) continue;
        float m = std::fmax(std::fmax(x.intensity[0], x.intensity[1]), x.intensity[2]);
        if (!(m * ratio512 < 0.005f)) continue;
        CubeLight o;
        std::memcpy(o.o, x.origin, sizeof o.o);
        std::memcpy(o.n, x.normal, sizeof o.n);
        std::memcpy(o.i, x.intensity, sizeof o.i);
        o.radius = x.radius;
        flagged.push_back(o);
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (!c->ambient) c->ambient = new AmbientData();
    AmbientData *D = c->ambient;
    D->nfaces = d->nfaces;
    D->lump_bytes = lump_bytes;
    D->has_sky = d->has_sky_ambient != 0;
    std::memcpy(D->sky, d->sky_ambient, sizeof D->sky);
    D->nflag = (int32_t)flagged.size();
    int rc;
    if ((rc = D->lights.ensure(std::max<size_t>(flagged.size(), 1) * sizeof(CubeLight), "cube lights")) ||
        (rc = D->faces.ensure(fd.size() * sizeof(AmbientFaceDev), "ambient faces")) ||
        (rc = D->avg.ensure(fd.size() * sizeof(float3), "face averages")) ||
        (rc = D->lump.ensure((size_t)std::max<int64_t>(lump_bytes, 4), "ambient lump")))
        return rc;
    VH_HIP(hipMemcpyAsync(D->faces.ptr, fd.data(), fd.size() * sizeof(AmbientFaceDev),
                          hipMemcpyHostToDevice, st));
    if (!flagged.empty())
        VH_HIP(hipMemcpyAsync(D->lights.ptr, flagged.data(), flagged.size() * sizeof(CubeLight),
                              hipMemcpyHostToDevice, st));
    if (lump_bytes > 0)
        VH_HIP(hipMemcpyAsync(D->lump.ptr, d->lump ? (const void *)d->lump : dev_lump, (size_t)lump_bytes,
                              d->lump ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
    if (d->nfaces > 0)
        hipLaunchKernelGGL(k_ambient_avg, dim3((uint32_t)((d->nfaces + 255) / 256)), dim3(256), 0, st,
                           (const AmbientFaceDev *)D->faces.ptr, d->nfaces, (const uint32_t *)D->lump.ptr,
                           (float3 *)D->avg.ptr);
    VH_HIP(hipGetLastError());
    VH_WAIT(c, st);
    return VH_OK;
}

int vh_ambient_cubes(vh_ctx *c, const float *points, int64_t n, float *cubes)
{
    clear_error();
    if (!c || n < 0 || (n > 0 && (!points || !cubes))) {
        set_error("bad leaf-ambient arguments");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->have_scene || !c->ambient) {
        set_error("no scene or no leaf-ambient inputs: call vh_scene_load and vh_ambient_load first");
        return VH_ESTATE;
    }
    if (n == 0) return VH_OK;
    if (n > INT32_MAX) {
        set_error("at most 2^31-1 points per call");
        return VH_EINVAL;
    }
    AmbientData *D = c->ambient;
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    int rc;
    if ((rc = D->points.ensure((size_t)n * 12, "ambient points")) ||
        (rc = D->cubes.ensure((size_t)n * 72, "ambient cubes")))
        return rc;
    VH_HIP(hipMemcpyAsync(D->points.ptr, points, (size_t)n * 12, hipMemcpyHostToDevice, st));
    AmbientArgs a;
    a.nodes = (const float4 *)c->nodes.ptr;
    a.tris = (const float4 *)c->tris.ptr;
    a.lights = (const CubeLight *)D->lights.ptr;
    a.nlights = D->nflag;
    a.faces = (const AmbientFaceDev *)D->faces.ptr;
    a.nfaces = D->nfaces;
    a.avg = (const float3 *)D->avg.ptr;
    a.lump = (const uint32_t *)D->lump.ptr;
    a.has_sky = D->has_sky;
    a.sky0 = D->sky[0];
    a.sky1 = D->sky[1];
    a.sky2 = D->sky[2];
    // DetMathF.Tan(7.275f * (MathF.PI / 180f)), correctly rounded
    a.tan_theta = (float)std::tan((double)(7.275f * (3.14159274f / 180.0f)));
    a.points = (const float *)D->points.ptr;
    a.npoints = n;
    a.cubes = (float *)D->cubes.ptr;
    {
        KernelTimer timer(c, VH_KERNEL_AMBIENT, st);
        hipLaunchKernelGGL(k_ambient, dim3((uint32_t)n), dim3(kFanBlock), 0, st, a);
        timer.done();
    }
    VH_HIP(hipGetLastError());
    VH_HIP(hipMemcpyAsync(cubes, D->cubes.ptr, (size_t)n * 72, hipMemcpyDeviceToHost, st));
    VH_WAIT(c, st);
    return VH_OK;
}

}  // extern "C"
