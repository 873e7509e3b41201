// kernels_tree.hpp -- device-resident tree topology: neighbour tables from
// the sorted Morton keys, and the regrid (refine closure, coarsen veto,
// order-preserving expand/contract with injection and restriction).
//
// The reference keeps the topology on the host and maintains neighbour
// records incrementally (include/ndtree/ndtree.hpp:727-1240,
// include/ndtree/neighbor.hpp:419-572), re-sorting and copying the flattened
// metadata to the device after every regrid (ndtree.hpp:1606-1700).  Here the
// leaves stay sorted on the device: refining a leaf replaces it in place by
// its children (their keys sort contiguously where the parent was) and
// coarsening a family replaces its contiguous run of children by the parent,
// so an exclusive scan over per-leaf output counts yields the new sorted order
// directly; neighbour records are then recomputed geometrically by binary
// search, which gives the same records as the reference's incremental
// bookkeeping (the oracle checks both against brute-force geometry).
//
// Per-item bodies are __host__ __device__ (run on the CPU by the emulation
// tests); the __global__ kernels distribute items over threads.
#pragma once

#include "device_common.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Summary written by the regrid kernels and read back by the host.
struct regrid_summary
{
    uint32_t total;     // new leaf count
    uint32_t n_refine;  // leaves refined (after closure)
    uint32_t n_coarsen; // families coarsened
    uint32_t err;       // bit 0: topology error, bit 1: root coarsening requested
    uint32_t max_level; // deepest level after the regrid
    uint32_t pad[3];
};

// Whether a planned regrid changes the tree and may be carried out: it
// refines or coarsens something, asks for no root coarsening and fits the
// capacity.  Read by the emit and the table build on the device, so that
// the host learns the outcome from one read-back after both.
__device__ inline bool regrid_applies(const regrid_summary* summary, uint32_t cap)
{
    const regrid_summary s = *summary;
    return !(s.err & 2u) && s.total <= cap && (s.n_refine | s.n_coarsen) != 0;
}

// Index of the leaf containing integer point p (resolution 2^depth) among the
// sorted keys, or -1 when no leaf contains it.
template <int R>
__host__ __device__ inline int32_t locate(const uint64_t* __restrict__ keys, uint32_t n, const uint32_t* p, int depth)
{
    const uint64_t mp     = interleave<R>(p);
    const uint64_t target = (mp << 6) | 63u;
    uint32_t       lo = 0, hi = n; // first index with key > target
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] <= target)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == 0) return -1;
    const uint64_t k    = keys[lo - 1];
    const uint64_t m    = k >> 6;
    const int      l    = key_level(k);
    const uint64_t span = uint64_t{ 1 } << (R * (depth - l));
    return (mp >= m && mp < m + span) ? static_cast<int32_t>(lo - 1) : -1;
}

// Neighbour record of (leaf, dir) by geometric search; false if the leaves do
// not tile the domain around it or break 2:1 face balance.
template <int R>
__host__ __device__ inline bool table_entry(const uint64_t* __restrict__ keys, uint32_t n, int depth, uint32_t leaf,
                                            int dir, face_link* out_link, uint32_t* out_meta)
{
    const uint64_t key = keys[leaf];
    const int      lv  = key_level(key);
    uint32_t       c[R];
    deinterleave<R>(key >> 6, c);
    const uint32_t mask = (1u << depth) - 1u;
    const uint32_t s    = 1u << (depth - lv);
    const int      dim  = dir >> 1;
    const bool     pos  = (dir & 1) != 0;
    const int      a    = R - 1 - dim; // physical axis of the normal
    uint32_t       p[R];
    for (int i = 0; i < R; ++i) p[i] = c[i];
    p[a]           = pos ? ((c[a] + s) & mask) : ((c[a] - 1u) & mask);
    face_link L    = { { -1, -1, -1, -1 } };
    uint32_t  meta = REL_NONE;
    bool      ok   = true;
    const int32_t j = locate<R>(keys, n, p, depth);
    if (j < 0)
        ok = false;
    else
    {
        const int lq = key_level(keys[j]);
        if (lq == lv)
        {
            L.ids[0] = j;
            meta     = REL_SAME;
        }
        else if (lq == lv - 1)
        {
            L.ids[0] = j;
            meta     = REL_COARSER;
            for (int jd = 0; jd < R; ++jd)
            {
                const uint32_t q = (jd == dim) ? (pos ? 0u : 1u) : ((c[R - 1 - jd] >> (depth - lv)) & 1u);
                meta |= q << (2 + jd);
            }
        }
        else if (lq == lv + 1)
        {
            meta = REL_FINER;
            for (int q = 0; q < (1 << (R - 1)); ++q)
            {
                uint32_t p2[R];
                for (int i = 0; i < R; ++i) p2[i] = c[i];
                p2[a]   = pos ? ((c[a] + s) & mask) : ((c[a] - s / 2) & mask);
                int bit = 0;
                for (int jd = 0; jd < R; ++jd)
                {
                    if (jd == dim) continue;
                    p2[R - 1 - jd] += ((q >> bit) & 1) * (s / 2);
                    ++bit;
                }
                const int32_t jj = locate<R>(keys, n, p2, depth);
                if (jj < 0 || key_level(keys[jj]) != lv + 1) ok = false;
                L.ids[q] = jj;
            }
        }
        else
            ok = false;
    }
    *out_link = L;
    *out_meta = meta;
    return ok;
}

// Refine mark: flag Refine and level < depth (ndtree.hpp:1887-1896).
__host__ __device__ inline uint8_t refine_mark(int8_t flag, uint64_t key, int depth)
{
    return (flag == 1 && key_level(key) < depth) ? 1 : 0;
}

// 2:1 refinement closure step for leaf i at `level` (ndtree.hpp:1129-1165):
// a refined leaf forces its coarser face neighbours to refine.
template <int R>
__host__ __device__ inline void close_item(const uint64_t* __restrict__ keys, int level,
                                           const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                                           uint8_t* mark, uint32_t i)
{
    if (!mark[i] || key_level(keys[i]) != level) return;
    for (int k = 0; k < 2 * R; ++k)
    {
        const size_t e = static_cast<size_t>(i) * 2 * R + k;
        if (rel_of(metas[e]) == REL_COARSER) mark[links[e].ids[0]] = 1;
    }
}

// Morton offset of a key within its parent (morton_id.hpp:55-62).
template <int R>
__host__ __device__ inline int offset_in_parent(uint64_t key, int depth)
{
    const int lv = key_level(key);
    if (lv == 0) return 0;
    uint32_t c[R];
    deinterleave<R>(key >> 6, c);
    int o = 0;
    for (int a = 0; a < R; ++a) o |= static_cast<int>((c[a] >> (depth - lv)) & 1u) << a;
    return o;
}

// Coarsening family test for leaf i (ndtree.hpp:904-940 eligibility,
// 1173-1239 veto): true when i is the first child of a complete family whose
// 2^R leaves are all flagged Coarsen and no boundary child has a finer face
// neighbour or a same-level face neighbour that refines.  *root_family is set
// when such a family would recombine into the root.
template <int R>
__host__ __device__ inline bool coarsen_family_head(const int8_t* __restrict__ flags, const uint64_t* __restrict__ keys,
                                                    uint32_t n, int depth, const face_link* __restrict__ links,
                                                    const uint32_t* __restrict__ metas,
                                                    const uint8_t* __restrict__ mark, uint32_t i, bool* root_family)
{
    constexpr int  F   = 1 << R;
    const uint64_t key = keys[i];
    const int      lv  = key_level(key);
    *root_family       = false;
    if (lv == 0 || offset_in_parent<R>(key, depth) != 0 || i + F > n) return false;
    uint32_t c[R];
    deinterleave<R>(key >> 6, c);
    const uint32_t s = 1u << (depth - lv);
    for (int o = 0; o < F; ++o)
    {
        uint32_t cc[R];
        for (int a = 0; a < R; ++a) cc[a] = c[a] + ((o >> a) & 1) * s;
        if (keys[i + o] != make_key<R>(cc, lv)) return false; // a child is not a leaf
        if (flags[i + o] != 2) return false;
    }
    if (lv == 1)
    {
        *root_family = true; // the reference aborts when the root's children recombine
        return false;
    }
    for (int o = 0; o < F; ++o)
    {
        if (mark[i + o]) return false;
        for (int k = 0; k < 2 * R; ++k)
        {
            const int  dim = k >> 1;
            const bool pos = (k & 1) != 0;
            const int  bit = (o >> (R - 1 - dim)) & 1; // child position on layout dim
            if (bit != (pos ? 1 : 0)) continue;       // not on that face of the parent
            const size_t   e   = static_cast<size_t>(i + o) * 2 * R + k;
            const uint32_t rel = rel_of(metas[e]);
            if (rel == REL_FINER) return false;
            if (rel == REL_SAME && mark[links[e].ids[0]]) return false;
        }
    }
    return true;
}

// Output leaves of old leaf i: refine -> 2^R, family head -> 1 (the parent),
// other family members -> 0, everything else -> 1.
template <int R>
__host__ __device__ inline uint32_t emit_count(const uint64_t* __restrict__ keys, int depth,
                                               const uint8_t* __restrict__ mark, const uint8_t* __restrict__ head,
                                               uint32_t i)
{
    const int o = offset_in_parent<R>(keys[i], depth);
    uint32_t  c = mark[i] ? (1u << R) : 1u;
    if (key_level(keys[i]) > 0 && static_cast<uint32_t>(o) <= i && head[i - o]) c = (o == 0) ? 1u : 0u;
    return c;
}

// Key of output j (0 <= j < count) of old leaf i.
template <int R>
__host__ __device__ inline uint64_t emit_key(uint64_t key, int depth, bool refine, bool family_head, int j)
{
    const int lv = key_level(key);
    uint32_t  c[R];
    deinterleave<R>(key >> 6, c);
    if (refine)
    {
        const uint32_t s = 1u << (depth - lv - 1);
        for (int a = 0; a < R; ++a) c[a] += ((j >> a) & 1) * s;
        return make_key<R>(c, lv + 1);
    }
    if (family_head)
    {
        const uint32_t s = 1u << (depth - lv);
        for (int a = 0; a < R; ++a) c[a] &= ~s;
        return make_key<R>(c, lv - 1);
    }
    return key;
}

// Cell t of child o of a refined leaf: injection (ndtree.hpp:1463-1497,
// 1528-1555).
template <int R>
__host__ __device__ inline double refine_value(box<R> const& B, const double* __restrict__ src, int o, int t)
{
    int tc[R], sc[R];
    B.coords(t, tc);
    for (int d = 0; d < R; ++d) sc[d] = ((o >> (R - 1 - d)) & 1) * B.half(d) + tc[d] / 2;
    return src[B.lin(sc)];
}

template <int R, int N>
__host__ __device__ inline double refine_value(const double* __restrict__ src, int o, int t)
{
    constexpr box<R> B = box<R>::cube(N);
    return refine_value<R>(B, src, o, t);
}

// Cell t of the parent of the family whose first child is slot i: mean of
// 2^R child cells summed from 0.0 in row-major order (ndtree.hpp:1499-1526,
// intergrid_operator.hpp:91-106).
template <int R>
__host__ __device__ inline double restrict_value(box<R> const& B, const double* __restrict__ u, uint32_t i, int t)
{
    int tc[R], base[R], kid = 0;
    B.coords(t, tc);
    for (int d = 0; d < R; ++d)
    {
        kid |= (tc[d] / B.half(d)) << (R - 1 - d);
        base[d] = (tc[d] % B.half(d)) * 2;
    }
    const double* q   = u + static_cast<size_t>(i + kid) * B.ni();
    double        sum = 0.0;
    for (int off = 0; off < box<R>::F; ++off)
    {
        int cc[R];
        for (int d = 0; d < R; ++d) cc[d] = base[d] + ((off >> (R - 1 - d)) & 1);
        sum += q[B.lin(cc)];
    }
    return sum / static_cast<double>(box<R>::F);
}

template <int R, int N>
__host__ __device__ inline double restrict_value(const double* __restrict__ u, uint32_t i, int t)
{
    constexpr box<R> B = box<R>::cube(N);
    return restrict_value<R>(B, u, i, t);
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <int R>
__global__ void k_build_table(const uint64_t* __restrict__ keys, uint32_t n, int depth, face_link* __restrict__ links,
                              uint32_t* __restrict__ metas, regrid_summary* __restrict__ summary, uint32_t guard_cap)
{
    // after a regrid planned on the device (guard_cap = the capacity): the
    // leaf count is the plan's, and nothing is built when the regrid leaves
    // the tree alone or stops with an error (see regrid_applies)
    if (guard_cap)
    {
        if (!regrid_applies(summary, guard_cap)) return;
        n = summary->total;
    }
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 2 * R) return;
    face_link L;
    uint32_t  meta;
    if (!table_entry<R>(keys, n, depth, idx / (2 * R), static_cast<int>(idx % (2 * R)), &L, &meta))
        atomicOr(&summary->err, 1u);
    links[idx] = L;
    metas[idx] = meta;
}

__global__ void k_mark_refine(const int8_t* __restrict__ flags, const uint64_t* __restrict__ keys, uint32_t n,
                              int depth, uint8_t* __restrict__ mark)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mark[i] = refine_mark(flags[i], keys[i], depth);
}

// Marks only flow to the next coarser level, so one pass per level from the
// finest down reaches the closure's fixed point.
template <int R>
__global__ void k_close_level(const uint64_t* __restrict__ keys, uint32_t n, int level,
                              const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                              uint8_t* __restrict__ mark)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) close_item<R>(keys, level, links, metas, mark, i);
}

template <int R>
__global__ void k_coarsen_families(const int8_t* __restrict__ flags, const uint64_t* __restrict__ keys, uint32_t n,
                                   int depth, const face_link* __restrict__ links,
                                   const uint32_t* __restrict__ metas, const uint8_t* __restrict__ mark,
                                   uint8_t* __restrict__ head, regrid_summary* __restrict__ summary)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool       root = false;
    const bool h    = coarsen_family_head<R>(flags, keys, n, depth, links, metas, mark, i, &root);
    head[i]         = h ? 1 : 0;
    if (root) atomicOr(&summary->err, 2u);
    if (h) atomicAdd(&summary->n_coarsen, 1u);
}

template <int R>
__global__ void k_emit_counts(const uint64_t* __restrict__ keys, uint32_t n, int depth,
                              const uint8_t* __restrict__ mark, const uint8_t* __restrict__ head,
                              uint32_t* __restrict__ counts, regrid_summary* __restrict__ summary)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    counts[i] = emit_count<R>(keys, depth, mark, head, i);
    if (mark[i]) atomicAdd(&summary->n_refine, 1u);
}

// Runs along +x (layout dim R-1; directions 2(R-1) = -x, 2(R-1)+1 = +x).  A
// leaf heads a run when its x index at its own level is a multiple of K or
// when its -x neighbour is not a same-level leaf; the run continues through
// same-level +x neighbours that are not heads, so it holds at most K leaves
// and every leaf belongs to exactly one run (a periodic same-level cycle
// always contains x index 0).
template <int R>
__host__ __device__ inline bool run_head(const uint64_t* keys, int depth, const uint32_t* metas, uint32_t i, uint32_t K)
{
    uint32_t c[R];
    deinterleave<R>(keys[i] >> 6, c);
    const uint32_t xl = c[0] >> (depth - key_level(keys[i]));
    return (xl % K) == 0 || rel_of(metas[static_cast<size_t>(i) * 2 * R + 2 * (R - 1)]) != REL_SAME;
}

template <int R>
__host__ __device__ inline uint32_t run_length(const uint64_t* keys, int depth, const face_link* links,
                                               const uint32_t* metas, uint32_t head, uint32_t K)
{
    uint32_t j = head, len = 1;
    while (len < K)
    {
        const size_t e = static_cast<size_t>(j) * 2 * R + 2 * (R - 1) + 1;
        if (rel_of(metas[e]) != REL_SAME) break;
        const uint32_t nb = static_cast<uint32_t>(links[e].ids[0]);
        if (run_head<R>(keys, depth, metas, nb, K)) break;
        j = nb;
        ++len;
    }
    return len;
}

template <int R>
__global__ void k_run_heads(const uint64_t* __restrict__ keys, uint32_t n, int depth,
                            const uint32_t* __restrict__ metas, uint32_t K, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = run_head<R>(keys, depth, metas, i, K) ? 1u : 0u;
}

template <int R>
__global__ void k_run_emit(const uint64_t* __restrict__ keys, uint32_t n, int depth,
                           const face_link* __restrict__ links, const uint32_t* __restrict__ metas, uint32_t K,
                           const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                           leaf_run* __restrict__ runs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) runs[pos[i]] = leaf_run{ i, run_length<R>(keys, depth, links, metas, i, K) };
}

// Exclusive scan of uint32 counts: per-block scan, scan of block sums in one
// workgroup, then add.  1024 items per block (256 threads x 4).
constexpr int kScanThreads = 256;
constexpr int kScanItems   = 4;
constexpr int kScanBlock   = kScanThreads * kScanItems;

template <int NT = kScanThreads>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[NT / 64];
    const int           lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t            x    = v;
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int j = 0; j < NT / 64; ++j)
    {
        if (j < w) off += wsum[j];
        tot += wsum[j];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// The regrid's planning passes for a tree of at most kPlanMaxLeaves leaves in
// one workgroup: refine marks, the 2:1 closure level by level from
// `top_level` down, the coarsening families, the output counts and their
// exclusive scan.  The per-leaf functions are those of k_mark_refine,
// k_close_level, k_coarsen_families, k_emit_counts and the scan; workgroup
// barriers take the place of the launches between the passes, so a small
// tree's regrid costs one launch instead of about a dozen.
constexpr int kPlanThreads   = 1024;
constexpr int kPlanMaxLeaves = 4 * kPlanThreads;

template <int R>
__global__ __launch_bounds__(kPlanThreads) void k_regrid_plan(
    const int8_t* __restrict__ flags, const uint64_t* __restrict__ keys, uint32_t n, int depth, int top_level,
    const face_link* __restrict__ links, const uint32_t* __restrict__ metas, uint8_t* mark, uint8_t* head,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ pos, regrid_summary* __restrict__ summary)
{
    const uint32_t t = threadIdx.x;
    if (t == 0) *summary = regrid_summary{}; // the counts below and the emit's deepest level start from 0
    for (uint32_t i = t; i < n; i += kPlanThreads) mark[i] = refine_mark(flags[i], keys[i], depth);
    __syncthreads();
    for (int l = top_level; l >= 1; --l)
    {
        for (uint32_t i = t; i < n; i += kPlanThreads) close_item<R>(keys, l, links, metas, mark, i);
        __syncthreads();
    }
    uint32_t n_coarsen = 0, n_refine = 0;
    bool     root_any  = false;
    for (uint32_t i = t; i < n; i += kPlanThreads)
    {
        bool       root = false;
        const bool h    = coarsen_family_head<R>(flags, keys, n, depth, links, metas, mark, i, &root);
        head[i]         = h ? 1 : 0;
        n_coarsen += h ? 1u : 0u;
        root_any = root_any || root;
    }
    __syncthreads();
    // counts over one contiguous chunk per thread, then the exclusive scan
    const uint32_t per = (n + kPlanThreads - 1) / kPlanThreads;
    const uint32_t lo  = t * per < n ? t * per : n;
    const uint32_t hi  = lo + per < n ? lo + per : n;
    uint32_t       sum = 0;
    for (uint32_t i = lo; i < hi; ++i)
    {
        const uint32_t c = emit_count<R>(keys, depth, mark, head, i);
        counts[i]        = c;
        sum += c;
        n_refine += mark[i] ? 1u : 0u;
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan<kPlanThreads>(sum, &total);
    for (uint32_t i = lo; i < hi; ++i)
    {
        pos[i] = off;
        off += counts[i];
    }
    if (n_coarsen) atomicAdd(&summary->n_coarsen, n_coarsen);
    if (n_refine) atomicAdd(&summary->n_refine, n_refine);
    if (root_any) atomicOr(&summary->err, 2u);
    if (t == 0) summary->total = total;
}

__global__ void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                              uint32_t* __restrict__ block_sums)
{
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t       v[kScanItems], s = 0;
    for (int j = 0; j < kScanItems; ++j)
    {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot);
    for (int j = 0; j < kScanItems; ++j)
    {
        if (base + j < n) out[base + j] = off;
        off += v[j];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(uint32_t* __restrict__ block_sums, uint32_t nblocks, regrid_summary* __restrict__ summary)
{
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kScanThreads)
    {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = (b < nblocks) ? block_sums[b] : 0u;
        uint32_t       tot;
        const uint32_t off = block_exclusive_scan(v, &tot);
        if (b < nblocks) block_sums[b] = carry + off;
        carry += tot;
    }
    if (threadIdx.x == 0) summary->total = carry;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ block_sums)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_sums[i / kScanBlock];
}

// New keys and interiors of old leaf i (one workgroup; every thread calls
// it): its copy, its 2^R injected children or its family's restriction.
template <int R, int NT>
__device__ inline void emit_leaf(box<R> const& B, const uint64_t* __restrict__ keys, const double* __restrict__ u,
                                 uint32_t i, int depth, const uint8_t* __restrict__ mark,
                                 const uint8_t* __restrict__ head, const uint32_t* __restrict__ counts,
                                 const uint32_t* __restrict__ pos, uint64_t* __restrict__ keys_out,
                                 double* __restrict__ u_out, regrid_summary* __restrict__ summary, int nfields,
                                 size_t fstride)
{
    const int      ni  = B.ni();
    const uint32_t cnt = counts[i];
    if (cnt == 0) return;
    const uint64_t key    = keys[i];
    const bool     refine = mark[i] != 0;
    const bool     fam    = head[i] != 0;
    const uint32_t dst    = pos[i];
    for (uint32_t j = 0; j < cnt; ++j)
    {
        // every field of the leaf goes through the same transfer
        for (int f = 0; f < nfields; ++f)
        {
            const double* uf  = u + f * fstride;
            double*       out = u_out + f * fstride + static_cast<size_t>(dst + j) * ni;
            if (refine)
            {
                const double* src = uf + static_cast<size_t>(i) * ni;
                for (int t = threadIdx.x; t < ni; t += NT) out[t] = refine_value<R>(B, src, static_cast<int>(j), t);
            }
            else if (fam)
            {
                for (int t = threadIdx.x; t < ni; t += NT) out[t] = restrict_value<R>(B, uf, i, t);
            }
            else
            {
                const double* src = uf + static_cast<size_t>(i) * ni;
                for (int t = threadIdx.x; t < ni; t += NT) out[t] = src[t];
            }
        }
        if (threadIdx.x == 0)
        {
            const uint64_t k  = emit_key<R>(key, depth, refine, fam, static_cast<int>(j));
            keys_out[dst + j] = k;
            // the maximum only grows: most leaves see it already reached and
            // skip the atomic (a stale read only costs an extra atomic)
            const uint32_t lv = static_cast<uint32_t>(key_level(k));
            if (lv > *reinterpret_cast<volatile uint32_t*>(&summary->max_level)) atomicMax(&summary->max_level, lv);
        }
    }
}

// New keys and interiors, one workgroup per old leaf.
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_emit(const uint64_t* __restrict__ keys, const double* __restrict__ u,
                                             uint32_t n, int depth, const uint8_t* __restrict__ mark,
                                             const uint8_t* __restrict__ head, const uint32_t* __restrict__ counts,
                                             const uint32_t* __restrict__ pos, uint64_t* __restrict__ keys_out,
                                             double* __restrict__ u_out, regrid_summary* __restrict__ summary,
                                             int nfields, size_t fstride, uint32_t guard_cap)
{
    constexpr box<R> B = box<R>::cube(N);
    if (guard_cap && !regrid_applies(summary, guard_cap)) return;
    const uint32_t i = xcd_item(blockIdx.x, gridDim.x);
    if (i >= n) return;
    emit_leaf<R, NT>(B, keys, u, i, depth, mark, head, counts, pos, keys_out, u_out, summary, nfields, fstride);
}

// k_emit for patches of any shape.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_emit_box(box<R> B, const uint64_t* __restrict__ keys,
                                                 const double* __restrict__ u, uint32_t n, int depth,
                                                 const uint8_t* __restrict__ mark, const uint8_t* __restrict__ head,
                                                 const uint32_t* __restrict__ counts,
                                                 const uint32_t* __restrict__ pos, uint64_t* __restrict__ keys_out,
                                                 double* __restrict__ u_out, regrid_summary* __restrict__ summary,
                                                 int nfields, size_t fstride, uint32_t guard_cap)
{
    if (guard_cap && !regrid_applies(summary, guard_cap)) return;
    const uint32_t i = xcd_item(blockIdx.x, gridDim.x);
    if (i >= n) return;
    emit_leaf<R, NT>(B, keys, u, i, depth, mark, head, counts, pos, keys_out, u_out, summary, nfields, fstride);
}

} // namespace amrhip
