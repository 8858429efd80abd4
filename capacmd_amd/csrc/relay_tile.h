// Tile geometry of the FP32 relay's TILES layout, shared by the assembly
// (cpim.hip), the product (solver.hip) and the dense expansion.
//
// The metal atoms are cut into Nb = ceil(Nm / 64) blocks of 64.  Tile t
// covers the block pair (I, J), I <= J, numbered row by row over the upper
// triangle: t = I Nb - I (I - 1) / 2 + (J - I).  Inside a tile, slot (k, l)
// (k = 0..63 the diagonal, l = 0..63 the lane) holds the pair
//     i = I*64 + l,   j = J*64 + ((l + k) & 63)
// as one float4 (g_a, g_b, c3, cq) at index t*kTileSlots + k*64 + l: at a fixed k
// the 64 lanes of a wavefront read 64 consecutive float4, each for its own
// row i, and the columns j they touch are a rotation of the block.  The
// relay is symmetric and its four coefficients depend on |r| only, so one
// slot serves both (i, j) and (j, i).  A diagonal tile (I == J) holds every
// ordered pair of its block (k = 1..63; k = 0 is i = j and holds zeros) and
// serves rows only: folding it to k <= 32 would save 3% of the bytes at
// Au2057 and cost the product a second code path.  Pairs with an atom past
// Nm hold zeros.
#pragma once

#include "capac_internal.h"

constexpr int kTileEdge = 64;                 // metal atoms per block (one per lane)
constexpr int kTileSlots = kTileEdge * kTileEdge;

__host__ __device__ __forceinline__ int tile_blocks(int Nm) { return (Nm + kTileEdge - 1) / kTileEdge; }

__host__ __device__ __forceinline__ long tile_count(int Nb) { return (long)Nb * (Nb + 1) / 2; }

__host__ __device__ __forceinline__ int tile_index(int I, int J, int Nb) {
    return I * Nb - I * (I - 1) / 2 + (J - I);
}

// (I, J) of tile t: walks the rows of the upper triangle (at most Nb steps)
__host__ __device__ __forceinline__ void tile_coords(int t, int Nb, int& I, int& J) {
    int i = 0, start = 0;
    while (start + (Nb - i) <= t) {
        start += Nb - i;
        ++i;
    }
    I = i;
    J = i + (t - start);
}

// the float4 slot that holds the pair (i, j), i != j
__host__ __device__ __forceinline__ long tile_slot(int i, int j, int Nb) {
    const int bi = i / kTileEdge, bj = j / kTileEdge, il = i % kTileEdge, jl = j % kTileEdge;
    int I, J, l, k;
    if (bi < bj) {
        I = bi; J = bj; l = il; k = (jl - il) & (kTileEdge - 1);
    } else if (bi > bj) {
        I = bj; J = bi; l = jl; k = (il - jl) & (kTileEdge - 1);
    } else {
        I = J = bi; l = il; k = (jl - il) & (kTileEdge - 1);
    }
    return (long)tile_index(I, J, Nb) * kTileSlots + (long)k * kTileEdge + l;
}
