// Trilinear corner weights and offsets of the indirect grid sample (IndirectGridSample3D.cu:432-572), shared by the
// grid-sample kernels and the regulariser's fused edge loss. Compiled without contraction (-ffp-contract=off), so
// every user rounds the weights identically.
#pragma once
#include "common.hpp"

namespace neat
{

struct GridArgs
{
    const float* grid;
    int64_t s0, s1, s2, s3, s4;  // element strides of (N, C, Z, Y, X)
    int64_t N;
    int32_t C, Z, Y, X;
};

struct Tri
{
    float w[8];
    int64_t off[8];  // element offset of each clamped corner (without the channel term)
    float ix, iy, iz;
    int32_t x0, y0, z0;
};

__device__ __forceinline__ int clip_i(int v, int lim) { return min(lim - 1, max(v, 0)); }

// The same weights with the corner offsets relative to the slot's first element, in 32 bits (the caller checks that
// a slot's extent fits): for kernels that sample one slot per corner set and want cheap address arithmetic.
struct TriLocal
{
    float w[8];
    int32_t off[8];
};

__device__ __forceinline__ TriLocal trilinear_local(const GridArgs& g, float u, float v, float wz)
{
    TriLocal t;
    float ix = ((u + 1) / 2) * (float)(g.X - 1);
    float iy = ((v + 1) / 2) * (float)(g.Y - 1);
    float iz = ((wz + 1) / 2) * (float)(g.Z - 1);
    int x0 = (int)floorf(ix), y0 = (int)floorf(iy), z0 = (int)floorf(iz);
    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    t.w[0] = ((float)x1 - ix) * ((float)y1 - iy) * ((float)z1 - iz);
    t.w[1] = (ix - (float)x0) * ((float)y1 - iy) * ((float)z1 - iz);
    t.w[2] = ((float)x1 - ix) * (iy - (float)y0) * ((float)z1 - iz);
    t.w[3] = (ix - (float)x0) * (iy - (float)y0) * ((float)z1 - iz);
    t.w[4] = ((float)x1 - ix) * ((float)y1 - iy) * (iz - (float)z0);
    t.w[5] = (ix - (float)x0) * ((float)y1 - iy) * (iz - (float)z0);
    t.w[6] = ((float)x1 - ix) * (iy - (float)y0) * (iz - (float)z0);
    t.w[7] = (ix - (float)x0) * (iy - (float)y0) * (iz - (float)z0);
    const int32_t ox0 = clip_i(x0, g.X) * (int32_t)g.s4, ox1 = clip_i(x1, g.X) * (int32_t)g.s4;
    const int32_t oy0 = clip_i(y0, g.Y) * (int32_t)g.s3, oy1 = clip_i(y1, g.Y) * (int32_t)g.s3;
    const int32_t oz0 = clip_i(z0, g.Z) * (int32_t)g.s2, oz1 = clip_i(z1, g.Z) * (int32_t)g.s2;
    t.off[0] = oz0 + oy0 + ox0;
    t.off[1] = oz0 + oy0 + ox1;
    t.off[2] = oz0 + oy1 + ox0;
    t.off[3] = oz0 + oy1 + ox1;
    t.off[4] = oz1 + oy0 + ox0;
    t.off[5] = oz1 + oy0 + ox1;
    t.off[6] = oz1 + oy1 + ox0;
    t.off[7] = oz1 + oy1 + ox1;
    return t;
}

__device__ __forceinline__ Tri trilinear(const GridArgs& g, int64_t n, float u, float v, float wz)
{
    Tri t;
    float ix = ((u + 1) / 2) * (float)(g.X - 1);
    float iy = ((v + 1) / 2) * (float)(g.Y - 1);
    float iz = ((wz + 1) / 2) * (float)(g.Z - 1);
    int x0 = (int)floorf(ix), y0 = (int)floorf(iy), z0 = (int)floorf(iz);
    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    t.w[0] = ((float)x1 - ix) * ((float)y1 - iy) * ((float)z1 - iz);
    t.w[1] = (ix - (float)x0) * ((float)y1 - iy) * ((float)z1 - iz);
    t.w[2] = ((float)x1 - ix) * (iy - (float)y0) * ((float)z1 - iz);
    t.w[3] = (ix - (float)x0) * (iy - (float)y0) * ((float)z1 - iz);
    t.w[4] = ((float)x1 - ix) * ((float)y1 - iy) * (iz - (float)z0);
    t.w[5] = (ix - (float)x0) * ((float)y1 - iy) * (iz - (float)z0);
    t.w[6] = ((float)x1 - ix) * (iy - (float)y0) * (iz - (float)z0);
    t.w[7] = (ix - (float)x0) * (iy - (float)y0) * (iz - (float)z0);
    int cx0 = clip_i(x0, g.X), cx1 = clip_i(x1, g.X);
    int cy0 = clip_i(y0, g.Y), cy1 = clip_i(y1, g.Y);
    int cz0 = clip_i(z0, g.Z), cz1 = clip_i(z1, g.Z);
    int64_t b = n * g.s0;
    t.off[0]  = b + cz0 * g.s2 + cy0 * g.s3 + cx0 * g.s4;
    t.off[1]  = b + cz0 * g.s2 + cy0 * g.s3 + cx1 * g.s4;
    t.off[2]  = b + cz0 * g.s2 + cy1 * g.s3 + cx0 * g.s4;
    t.off[3]  = b + cz0 * g.s2 + cy1 * g.s3 + cx1 * g.s4;
    t.off[4]  = b + cz1 * g.s2 + cy0 * g.s3 + cx0 * g.s4;
    t.off[5]  = b + cz1 * g.s2 + cy0 * g.s3 + cx1 * g.s4;
    t.off[6]  = b + cz1 * g.s2 + cy1 * g.s3 + cx0 * g.s4;
    t.off[7]  = b + cz1 * g.s2 + cy1 * g.s3 + cx1 * g.s4;
    t.ix      = ix;
    t.iy      = iy;
    t.iz      = iz;
    t.x0      = x0;
    t.y0      = y0;
    t.z0      = z0;
    return t;
}

}  // namespace neat
