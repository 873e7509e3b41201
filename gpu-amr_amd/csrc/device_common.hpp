// device_common.hpp -- shared device helpers: Morton keys, patch geometry,
// neighbour-table encoding and the XCD-aware workgroup mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#pragma clang fp contract(off)

namespace amrhip
{

constexpr int kMaxDepth  = 16; // keys hold (rank * depth + 6) <= 64 bits
constexpr int kMaxLevels = kMaxDepth + 1;
constexpr int kXcds      = 8;  // MI355X: 8 accelerator complex dies

// Neighbour relation codes (include/cuda/halo_exchange.hpp:11-17).
enum : uint32_t
{
    REL_NONE    = 0,
    REL_SAME    = 1,
    REL_FINER   = 2,
    REL_COARSER = 3,
};

// Device neighbour table, one entry per (leaf, direction):
//   ids[4]: same/coarser -> ids[0] = neighbour slot; finer -> the 2^(rank-1)
//           fine slots little-endian over the non-normal layout dims.
//   meta:   bits 0-1 relation, bit 2+d = contact quadrant of layout dim d.
struct face_link
{
    int32_t ids[4];
};

__host__ __device__ inline uint32_t rel_of(uint32_t meta) { return meta & 3u; }
__host__ __device__ inline uint32_t quad_of(uint32_t meta, int d) { return (meta >> (2 + d)) & 1u; }

// ---------------------------------------------------------------------------
// Morton keys: key = (interleave(x, y[, z]) << 6) | level, x in bit 0.
// ---------------------------------------------------------------------------
template <int R>
__host__ __device__ inline uint64_t interleave(const uint32_t* c)
{
    uint64_t m = 0;
    for (int b = 0; b < kMaxDepth; ++b)
        for (int i = 0; i < R; ++i) m |= static_cast<uint64_t>((c[i] >> b) & 1u) << (R * b + i);
    return m;
}

template <int R>
__host__ __device__ inline void deinterleave(uint64_t m, uint32_t* c)
{
    for (int i = 0; i < R; ++i) c[i] = 0;
    for (int b = 0; b < kMaxDepth; ++b)
        for (int i = 0; i < R; ++i) c[i] |= static_cast<uint32_t>((m >> (R * b + i)) & 1u) << b;
}

__host__ __device__ inline int key_level(uint64_t k) { return static_cast<int>(k & 0x3Fu); }

// A run: `len` same-level leaves that follow each other along +x, starting
// at leaf `head` (see run_head / run_length).  The run-streaming step kernels
// give each run to one wave, so the x faces inside a run move through
// registers instead of memory.
struct leaf_run
{
    uint32_t head, len;
};

template <int R>
__host__ __device__ inline uint64_t make_key(const uint32_t* c, int level)
{
    return (interleave<R>(c) << 6) | static_cast<uint64_t>(level & 0x3F);
}

// ---------------------------------------------------------------------------
// Compact patch layout: n^R interior doubles, row-major, last index fastest.
// Layout dim j holds physical axis R-1-j.
// ---------------------------------------------------------------------------
template <int R, int N>
struct patch_shape
{
    static constexpr int NI   = (R == 2) ? N * N : N * N * N; // interior cells
    static constexpr int FACE = (R == 2) ? N : N * N;         // cells per face
    static constexpr int F    = 1 << R;                       // children
    static constexpr int K    = 1 << (R - 1);                 // face children
    static constexpr int NDIR = 2 * R;
    static constexpr int HALF = N / 2;

    // compact interior index of layout coordinates c[0..R-1]
    __host__ __device__ static int lin(const int* c)
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l = l * N + c[d];
        return l;
    }
    __host__ __device__ static void coords(int l, int* c)
    {
        for (int d = R - 1; d >= 0; --d)
        {
            c[d] = l % N;
            l /= N;
        }
    }
};

// ---------------------------------------------------------------------------
// Patch extents of any shape: layout dim d holds n(d) interior cells (the
// reference's static_shape<N, M[, O]>: d = 0 slowest, the last is physical
// x), row-major with the last index fastest.  Every extent is even (a child
// covers half of each dim).  The general kernels take a box as an argument;
// the fixed-shape kernels build one from constants, which the compiler folds.
// ---------------------------------------------------------------------------
template <int R>
struct box
{
    static constexpr int F    = 1 << R;       // children
    static constexpr int K    = 1 << (R - 1); // face children
    static constexpr int NDIR = 2 * R;

    int32_t e[3];

    static constexpr box cube(int n) { return box{ { n, n, n } }; }

    __host__ __device__ constexpr int n(int d) const { return e[d]; }
    __host__ __device__ constexpr int half(int d) const { return e[d] / 2; }
    __host__ __device__ constexpr int ni() const { return (R == 2) ? e[0] * e[1] : e[0] * e[1] * e[2]; }
    // cells of a face normal to layout dim `dim`
    __host__ __device__ constexpr int face(int dim) const { return ni() / e[dim]; }
    // stride of layout dim d in the compact interior
    __host__ __device__ constexpr int stride(int d) const
    {
        int s = 1;
        for (int k = R - 1; k > d; --k) s *= e[k];
        return s;
    }
    __host__ __device__ constexpr int lin(const int* c) const
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l = l * e[d] + c[d];
        return l;
    }
    __host__ __device__ constexpr void coords(int l, int* c) const
    {
        for (int d = R - 1; d >= 0; --d)
        {
            c[d] = l % e[d];
            l /= e[d];
        }
    }
    // padded view with h ghost layers per side: (n(d) + 2h) per dim
    __host__ __device__ constexpr int padded(int h) const
    {
        int p = 1;
        for (int d = 0; d < R; ++d) p *= e[d] + 2 * h;
        return p;
    }
};

// ---------------------------------------------------------------------------
// Workgroup -> item mapping.  Workgroups are dispatched round-robin over the
// XCDs (workgroup b lands on XCD b % 8).  Giving each XCD one contiguous range
// of Morton-ordered leaves keeps face neighbours inside the same L2.  The grid
// must be a multiple of kXcds; items >= count are skipped by the caller.
// ---------------------------------------------------------------------------
__device__ inline uint32_t xcd_item(uint32_t b, uint32_t nblocks)
{
    const uint32_t per = nblocks / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

inline uint32_t xcd_grid(uint64_t items)
{
    const uint64_t per = (items + kXcds - 1) / kXcds;
    return static_cast<uint32_t>(per * kXcds);
}

} // namespace amrhip
