// cpu_ref.hpp -- CPU oracle for the patch-AMR finite-volume advection step.
//
// TEST INFRASTRUCTURE ONLY. This is a clean-room restatement of the
// reference's CPU algorithm (bgce-cse/gpu-amr) used as the checker for the
// HIP implementation and as the timed CPU baseline. Product code never links
// or calls it; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (through cpu_ref_capi.cpp).
//
// Every function cites the reference file:line it restates. Arithmetic is
// written operation by operation in the reference's order and the library is
// compiled with -ffp-contract=off, so results are reproducible bit for bit.
//
// Pinning (how this oracle is checked; the reference itself does not build
// in this image, see the oracle section of the design notes):
//   * Morton keys / parent / child / offset   -> closed-form bit interleave
//                                                  (tests/test_oracle_keys.py)
//   * neighbour tables after any regrid        -> independent brute-force
//                                                  geometric search
//                                                  (tests/brute.py)
//   * halo exchange on mixed trees             -> brute-force geometric halo
//                                                  fill (tests/brute.py)
//   * uniform periodic step                    -> dense N x N (x N) periodic
//                                                  upwind sweep (numpy)
//   * refine closure / coarsen veto            -> geometric restatement of the
//                                                  2:1 rule (tests/brute.py)
//   * restriction / injection                  -> closed forms
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace oracle
{

using key_t = std::uint64_t;

// ---------------------------------------------------------------------------
// Morton keys.  key = (interleave(x, y[, z]) << 6) | level
// (include/morton/morton_id.hpp:132-140 (2D), :350-359 (3D)); x occupies bit
// 0 of each interleaved group, y bit 1, z bit 2
// (lib/libmorton/morton2D.h:113-116, lib/libmorton/morton3D.h:119-121).
// ---------------------------------------------------------------------------
template <int R>
struct morton
{
    static_assert(R == 2 || R == 3);
    using coords_t = std::array<std::uint32_t, R>;

    int depth;

    static std::uint64_t interleave(coords_t const& c)
    {
        std::uint64_t m = 0;
        for (int b = 0; b < 21; ++b)
            for (int i = 0; i < R; ++i)
                m |= static_cast<std::uint64_t>((c[i] >> b) & 1u) << (R * b + i);
        return m;
    }

    static coords_t deinterleave(std::uint64_t m)
    {
        coords_t c{};
        for (int b = 0; b < 21; ++b)
            for (int i = 0; i < R; ++i)
                c[i] |= static_cast<std::uint32_t>((m >> (R * b + i)) & 1u) << b;
        return c;
    }

    // morton_id::encode (morton_id.hpp:132-140, :350-359)
    key_t encode(coords_t const& c, int level) const
    {
        return (interleave(c) << 6) | static_cast<key_t>(level & 0x3F);
    }

    // morton_id::decode (morton_id.hpp:118-130, :334-348)
    std::pair<coords_t, int> decode(key_t k) const
    {
        return { deinterleave(k >> 6), static_cast<int>(k & 0x3F) };
    }

    static int level(key_t k) { return static_cast<int>(k & 0x3F); }

    // morton_id::parent_of (morton_id.hpp:103-116, :319-332)
    key_t parent_of(key_t k) const
    {
        auto [c, l] = decode(k);
        if (l <= 0) throw std::logic_error("root has no parent");
        const std::uint32_t off = 1u << (depth - l);
        for (int i = 0; i < R; ++i) c[i] &= ~off;
        return encode(c, l - 1);
    }

    // morton_id::offset (morton_id.hpp:64-83, :277-298)
    key_t offset(key_t id, unsigned o) const
    {
        auto [c, l] = decode(id);
        const std::uint32_t delta = 1u << (depth - l);
        c[0] += (o & 1u) * delta;
        c[1] += ((o >> 1) & 1u) * delta;
        if constexpr (R == 3) c[2] += ((o >> 2) & 1u) * delta;
        return encode(c, l);
    }

    // morton_id::child_of (morton_id.hpp:142-148, :361-367): the child key is
    // built from parent_id + 1 (the level field incremented).
    key_t child_of(key_t parent, unsigned o) const
    {
        if (level(parent) >= depth) throw std::logic_error("cell already at max level");
        return offset(parent + 1, o);
    }

    // morton_id::offset_of (morton_id.hpp:55-62, :267-275)
    unsigned offset_of(key_t k) const
    {
        auto [c, l] = decode(k);
        const std::uint32_t delta = 1u << (depth - l + 1);
        unsigned o = 0;
        for (int i = 0; i < R; ++i) o += (c[i] % delta == 0 ? 0u : 1u) << i;
        return o;
    }
};

// ---------------------------------------------------------------------------
// Neighbour variant (include/ndtree/neighbor.hpp:22-100) holding patch keys.
// ---------------------------------------------------------------------------
enum class rel : std::int8_t
{
    none    = 0,
    same    = 1,
    finer   = 2,
    coarser = 3,
};

template <int R>
struct neighbor
{
    static constexpr int K = 1 << (R - 1); // finer::num_neighbors (neighbor.hpp:73-76)
    rel                  kind = rel::none;
    key_t                id   = 0;            // same / coarser
    std::array<key_t, K> ids{};               // finer (value-initialised: key 0)
    std::array<int, R>   quadrant{};          // coarser contact quadrant
};

// Reference device metadata record (include/cuda/halo_exchange.hpp:11-26);
// the flattened, linear-index form of one (leaf, direction) table entry.
struct halo_meta
{
    std::int32_t neighbor    = -1;
    std::int32_t finer_ids[4] = { -1, -1, -1, -1 };
    std::int32_t quadrant[3] = { 0, 0, 0 };
    std::int8_t  relation    = 0;
    std::int8_t  padding[3]  = { 0, 0, 0 };
};
static_assert(sizeof(halo_meta) == 36);

enum refine_status : std::int8_t
{
    Stable  = 0,
    Refine  = 1,
    Coarsen = 2,
};

// ---------------------------------------------------------------------------
// Simple static-partition parallel for-each over [0, n): the restatement of
// std::for_each(std::execution::par_unseq, ...) over patches
// (include/config/definitions.hpp:19-21, include/solver/amr_solver.hpp:274-277).
// ---------------------------------------------------------------------------
template <typename Fn>
inline void parallel_for(std::size_t n, int threads, Fn&& fn)
{
    if (threads <= 1 || n < 2)
    {
        for (std::size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    const std::size_t t = std::min<std::size_t>(static_cast<std::size_t>(threads), n);
    std::vector<std::thread> pool;
    pool.reserve(t);
    for (std::size_t w = 0; w < t; ++w)
    {
        pool.emplace_back(
            [&, w]
            {
                const std::size_t lo = n * w / t, hi = n * (w + 1) / t;
                for (std::size_t i = lo; i < hi; ++i) fn(i);
            }
        );
    }
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// The tree + solver restatement.  R = rank, H = halo width, Ns = interior
// cells of layout dims 0..R-1 (static_shape<N, M[, O]>).  Layout is row-major with the last
// index fastest (include/containers/static_layout.hpp:28-37); physical x is the
// last layout index (include/solver/physics_system.hpp:126-135,215-224).
// ---------------------------------------------------------------------------
// rusanovFlux of AdvectionPhysics (AdvectionPhysics.hpp:45-66) for velocity
// component v along the face normal: 0.5 (ul v + ur v) - 0.5 |v| (ur - ul).
inline double advection_flux(double ul, double ur, double v)
{
    const double fl   = ul * v;
    const double fr   = ur * v;
    const double smax = std::abs(v);
    return 0.5 * (fl + fr) - 0.5 * smax * (ur - ul);
}

template <int R, int H, int... Ns>
class tree
{
public:
    static_assert(sizeof...(Ns) == R, "one extent per layout dim");
    static constexpr std::array<int, R> NS{ Ns... };             // interior cells per layout dim
    static constexpr std::array<int, R> PS{ (Ns + 2 * H)... };   // padded extents
    static constexpr int FLAT = ((Ns + 2 * H) * ...);            // padded cells
    static constexpr int NI   = (Ns * ...);                      // interior cells
    static constexpr int F    = 1 << R;                          // nd fanout
    static constexpr int K    = 1 << (R - 1);                    // face children
    static constexpr int NDIR = 2 * R;
    static_assert(((Ns % 2 == 0 && Ns >= 2 * H) && ...), "even extents; a halo must not span more than one finer cell");

    using idx_t      = std::array<int, R>;
    using neighbor_t = neighbor<R>;
    using nbrs_t     = std::array<neighbor_t, NDIR>;

    static constexpr std::array<int, R> strides()
    {
        std::array<int, R> s{};
        s[R - 1] = 1;
        for (int d = R - 2; d >= 0; --d) s[d] = s[d + 1] * PS[d + 1];
        return s;
    }
    static constexpr std::array<int, R> S = strides();

    static int lin(idx_t const& i)
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l += i[d] * S[d];
        return l;
    }

    // Row-major loop over a box [lo, hi) (container_manipulations.hpp:75-123).
    template <typename Fn>
    static void box_for(idx_t const& lo, idx_t const& hi, Fn&& fn)
    {
        idx_t i = lo;
        if constexpr (R == 2)
        {
            for (i[0] = lo[0]; i[0] < hi[0]; ++i[0])
                for (i[1] = lo[1]; i[1] < hi[1]; ++i[1]) fn(i);
        }
        else
        {
            for (i[0] = lo[0]; i[0] < hi[0]; ++i[0])
                for (i[1] = lo[1]; i[1] < hi[1]; ++i[1])
                    for (i[2] = lo[2]; i[2] < hi[2]; ++i[2]) fn(i);
        }
    }

    template <typename Fn>
    static void interior_for(Fn&& fn)
    {
        idx_t lo, hi;
        for (int d = 0; d < R; ++d)
        {
            lo[d] = H;
            hi[d] = H + NS[d];
        }
        box_for(lo, hi, fn);
    }

    // Direction helpers (neighbor.hpp:156-211): dir k -> layout dim k/2,
    // negative iff k is even, opposite = k ^ 1.
    static int  dim_of(int k) { return k / 2; }
    static bool positive(int k) { return (k % 2) == 1; }
    static int  opposite(int k) { return (k % 2 == 0) ? k + 1 : k - 1; }

    // Child hypercube (fanout 2, rank R) row-major multi-index.
    static idx_t child_multi(int flat)
    {
        idx_t c{};
        for (int d = R - 1; d >= 0; --d)
        {
            c[d] = flat % 2;
            flat /= 2;
        }
        return c;
    }
    static int child_linear(idx_t const& c)
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l = l * 2 + c[d];
        return l;
    }

    // neighbor.hpp:291-313
    static bool is_parent_neighbor(int child, int k)
    {
        const auto c = child_multi(child);
        return positive(k) ? (c[dim_of(k)] == 1) : (c[dim_of(k)] == 0);
    }

    // neighbor.hpp:316-337
    static int fine_boundary_linear_index(idx_t const& c, int normal)
    {
        int l = 0, m = 1;
        for (int d = 0; d < R; ++d)
        {
            if (d == normal) continue;
            l += c[d] * m;
            m *= 2;
        }
        return l;
    }

    // neighbor.hpp:340-364
    static std::array<int, R> contact_quadrant(int idx, int k)
    {
        std::array<int, R> q{};
        for (int i = 0; i < R; ++i)
        {
            if (i == dim_of(k))
                q[i] = positive(k) ? 0 : 1;
            else
            {
                q[i] = idx % 2;
                idx /= 2;
            }
        }
        return q;
    }

    // neighbor.hpp:383-395
    static int sibling_offset(idx_t c, int k)
    {
        const int d = dim_of(k);
        c[d] = positive(k) ? (c[d] + 1) % 2 : (c[d] + 2 - 1) % 2;
        return child_linear(c);
    }

    // neighbor.hpp:397-417
    static std::array<int, K> boundary_children(int k)
    {
        std::array<int, K> bc{};
        for (int i = 0; i < F; ++i)
        {
            if (!is_parent_neighbor(i, k)) continue;
            bc[fine_boundary_linear_index(child_multi(i), dim_of(k))] = i;
        }
        return bc;
    }

public:
    tree(int depth, std::size_t capacity, std::array<double, 3> lengths)
        : m_morton{ depth }, m_capacity{ capacity }, m_lengths{ lengths }
    {
        m_keys.resize(capacity);
        m_status.resize(capacity, Stable);
        m_nbrs.resize(capacity);
        m_cur.assign(capacity * FLAT, 0.0);
        m_next.assign(capacity * FLAT, 0.0);
        // ndtree ctor: periodic root (ndtree.hpp:411-421)
        nbrs_t root{};
        for (int k = 0; k < NDIR; ++k)
        {
            root[k].kind = rel::same;
            root[k].id   = 0;
        }
        append(0, root);
    }

    std::size_t size() const { return m_size; }
    key_t       key_at(std::size_t i) const { return m_keys.at(i); }
    int         depth() const { return m_morton.depth; }
    morton<R> const& keys() const { return m_morton; }
    double*       patch(std::size_t i) { return &m_cur[i * FLAT]; }
    double const* patch(std::size_t i) const { return &m_cur[i * FLAT]; }
    double*       next_patch(std::size_t i) { return &m_next[i * FLAT]; }
    void          set_threads(int t) { m_threads = std::max(1, t); }
    void          set_velocity(std::array<double, 3> v) { m_vel = v; }
    void          set_cfl(double c) { m_cfl = c; }
    std::int8_t   refine_status_at(std::size_t i) const { return m_status.at(i); }
    std::vector<key_t> const& last_refine_set() const { return m_last_refine; }
    std::vector<key_t> const& last_coarsen_set() const { return m_last_coarsen; }

    // ---------------------------------------------------------------------
    // Geometry: physics_system (physics_system.hpp:150-231)
    // ---------------------------------------------------------------------
    std::array<double, R> patch_sizes(key_t k) const
    {
        const int                 l = morton<R>::level(k);
        const std::uint32_t       pm = 1u << (m_morton.depth - l);
        const double              maxc = static_cast<double>(1u << m_morton.depth);
        std::array<double, R>     r{};
        for (int i = 0; i < R; ++i) r[i] = m_lengths[i] * static_cast<double>(pm) / maxc;
        return r;
    }
    std::array<double, R> cell_sizes(key_t k) const
    {
        auto ps = patch_sizes(k);
        std::array<double, R> r{};
        for (int i = 0; i < R; ++i) r[i] = ps[i] / static_cast<double>(NS[R - 1 - i]); // physical axis i
        return r;
    }
    std::array<double, R> patch_coord(key_t k) const
    {
        auto [c, l]  = m_morton.decode(k);
        const double maxc = static_cast<double>(1u << m_morton.depth);
        std::array<double, R> r{};
        for (int i = 0; i < R; ++i) r[i] = m_lengths[i] * static_cast<double>(c[i]) / maxc;
        return r;
    }
    // cell origin of padded multi-index idx (physics_system.hpp:199-231);
    // physical axis i reads layout index R-1-i.
    std::array<double, R> cell_coord(key_t k, idx_t const& idx) const
    {
        auto o  = patch_coord(k);
        auto cs = cell_sizes(k);
        std::array<double, R> r{};
        for (int i = 0; i < R; ++i)
        {
            const double dc = static_cast<double>(idx[R - 1 - i]) - static_cast<double>(H);
            r[i]            = o[i] + dc * cs[i];
        }
        return r;
    }
    // initialize(): cell centre = origin + 0.5 * size (amr_solver.hpp:105-145).
    void cell_centers(std::size_t i, double* out) const
    {
        const key_t k  = m_keys[i];
        const auto  cs = cell_sizes(k);
        int         c  = 0;
        interior_for(
            [&](idx_t const& idx)
            {
                auto o = cell_coord(k, idx);
                for (int d = 0; d < R; ++d) out[c * R + d] = o[d] + 0.5 * cs[d];
                ++c;
            }
        );
    }

    // Interior values in row-major interior order.
    void set_interior(std::size_t i, double const* v)
    {
        int c = 0;
        interior_for([&](idx_t const& idx) { m_cur[i * FLAT + lin(idx)] = v[c++]; });
    }
    void get_interior(std::size_t i, double* v) const
    {
        int c = 0;
        interior_for([&](idx_t const& idx) { v[c++] = m_cur[i * FLAT + lin(idx)]; });
    }

    // ---------------------------------------------------------------------
    // Tree topology (ndtree.hpp)
    // ---------------------------------------------------------------------

    // ndtree::append (ndtree.hpp:1304-1316), with the capacity check the
    // reference lacks.
    void append(key_t k, nbrs_t const& n)
    {
        if (m_size >= m_capacity) throw std::length_error("tree capacity exceeded");
        m_keys[m_size] = k;
        m_map[k]       = m_size;
        m_nbrs[m_size] = n;
        ++m_size;
    }

    std::size_t map_at(key_t k) const
    {
        auto it = m_map.find(k);
        if (it == m_map.end()) throw std::out_of_range("patch key not in index map");
        return it->second;
    }

    // neighbor_utils::compute_child_neighbors (neighbor.hpp:419-493)
    nbrs_t compute_child_neighbors(key_t parent, nbrs_t const& pn, int child) const
    {
        const idx_t cm = child_multi(child);
        nbrs_t      out{};
        for (int k = 0; k < NDIR; ++k)
        {
            if (!is_parent_neighbor(child, k) || parent == 0)
            {
                out[k].kind = rel::same;
                out[k].id   = m_morton.child_of(parent, static_cast<unsigned>(sibling_offset(cm, k)));
                continue;
            }
            neighbor_t const& p = pn[k];
            switch (p.kind)
            {
                case rel::none: out[k].kind = rel::none; break;
                case rel::same:
                {
                    out[k].kind = rel::coarser;
                    out[k].id   = p.id;
                    // compute_neighbor_coarse_block_coords (neighbor.hpp:366-379)
                    for (int d = 0; d < R; ++d) out[k].quadrant[d] = cm[d];
                    out[k].quadrant[dim_of(k)] = positive(k) ? 0 : 1;
                    break;
                }
                case rel::finer:
                {
                    out[k].kind = rel::same;
                    out[k].id   = p.ids[fine_boundary_linear_index(cm, dim_of(k))];
                    break;
                }
                case rel::coarser: out[k].kind = rel::none; break;
            }
        }
        return out;
    }

    // neighbor_utils::compute_parent_neighbors (neighbor.hpp:495-572): the
    // parent's relation comes from the first boundary child (:505-508).
    nbrs_t compute_parent_neighbors(std::array<nbrs_t, F> const& cn) const
    {
        nbrs_t out{};
        for (int k = 0; k < NDIR; ++k)
        {
            const auto        bc    = boundary_children(k);
            neighbor_t const& first = cn[bc[0]][k];
            switch (first.kind)
            {
                case rel::none: out[k].kind = rel::none; break;
                case rel::same:
                {
                    out[k].kind = rel::finer;
                    out[k].ids  = {};
                    for (int i = 0; i < K; ++i)
                    {
                        neighbor_t const& c = cn[bc[i]][k];
                        if (c.kind == rel::same) out[k].ids[i] = c.id;
                    }
                    break;
                }
                case rel::coarser:
                    out[k].kind = rel::same;
                    out[k].id   = first.id;
                    break;
                case rel::finer:
                    throw std::logic_error("child has finer neighbor during coarsening");
            }
        }
        return out;
    }

    // ndtree::enforce_symmetry_after_refinement (ndtree.hpp:942-1037)
    void enforce_symmetry_after_refinement(key_t parent)
    {
        const nbrs_t pn = m_nbrs[map_at(parent)];
        for (int k = 0; k < NDIR; ++k)
        {
            neighbor_t const& n   = pn[k];
            const int         opp = opposite(k);
            const auto        bc  = boundary_children(k);
            if (n.kind == rel::same)
            {
                neighbor_t nn{};
                nn.kind = rel::finer;
                for (int i = 0; i < K; ++i)
                    nn.ids[i] = m_morton.child_of(parent, static_cast<unsigned>(bc[i]));
                m_nbrs[map_at(n.id)][opp] = nn;
            }
            else if (n.kind == rel::finer)
            {
                for (int i = 0; i < K; ++i)
                {
                    neighbor_t nn{};
                    nn.kind = rel::same;
                    nn.id   = m_morton.child_of(parent, static_cast<unsigned>(bc[i]));
                    m_nbrs[map_at(n.ids[i])][opp] = nn;
                }
            }
            else if (n.kind == rel::coarser)
            {
                throw std::logic_error("coarser neighbor during refinement");
            }
        }
    }

    // ndtree::enforce_symmetry_after_recombining (ndtree.hpp:1039-1125)
    void enforce_symmetry_after_recombining(key_t parent, nbrs_t const& pn)
    {
        for (int k = 0; k < NDIR; ++k)
        {
            const int         opp = opposite(k);
            neighbor_t const& n   = pn[k];
            if (n.kind == rel::same)
            {
                auto it = m_map.find(n.id);
                if (it != m_map.end())
                {
                    neighbor_t nn{};
                    nn.kind                 = rel::same;
                    nn.id                   = parent;
                    m_nbrs[it->second][opp] = nn;
                }
            }
            else if (n.kind == rel::finer)
            {
                for (int i = 0; i < K; ++i)
                {
                    neighbor_t nn{};
                    nn.kind     = rel::coarser;
                    nn.id       = parent;
                    nn.quadrant = contact_quadrant(i, opp);
                    m_nbrs[map_at(n.ids[i])][opp] = nn;
                }
            }
            else if (n.kind == rel::coarser)
            {
                throw std::logic_error("coarser neighbor after recombining");
            }
        }
    }

    // ndtree::patch_mapping_impl (ndtree.hpp:1463-1497): for each coarse
    // interior cell -> (child position, 2^R fine cells in row-major order).
    template <typename Fn>
    static void patch_mapping(Fn&& fn)
    {
        interior_for(
            [&](idx_t const& idxs)
            {
                idx_t fine_patch{}, base{};
                for (int d = 0; d < R; ++d)
                {
                    const int section = NS[d] / 2;
                    fine_patch[d]     = (idxs[d] - H) / section;
                    base[d]           = ((idxs[d] - H) % section) * 2 + H;
                }
                const int fp = child_linear(fine_patch);
                std::array<int, F> fl{};
                int                o = 0;
                idx_t              zero{}, two;
                two.fill(2);
                box_for(
                    zero, two,
                    [&](idx_t const& off)
                    {
                        idx_t c = base;
                        for (int d = 0; d < R; ++d) c[d] += off[d];
                        fl[o++] = lin(c);
                    }
                );
                fn(fp, fl, lin(idxs));
            }
        );
    }

    // ndtree::interpolate_patch + linear_interpolator::interpolation
    // (ndtree.hpp:1528-1555, intergrid_operator.hpp:40-73): injection.
    void interpolate_patch(std::size_t from, std::size_t start_to)
    {
        patch_mapping(
            [&](int fp, std::array<int, F> const& fl, int coarse)
            {
                double*       to  = &m_cur[(start_to + fp) * FLAT];
                double const* src = &m_cur[from * FLAT];
                for (int i = 0; i < F; ++i) to[fl[i]] = src[coarse];
            }
        );
    }

    // ndtree::restrict_patches + linear_interpolator::restriction
    // (ndtree.hpp:1499-1526, intergrid_operator.hpp:91-106): mean, summed in
    // row-major fine-cell order starting from 0.
    void restrict_patches(std::size_t start_from, std::size_t to)
    {
        patch_mapping(
            [&](int fp, std::array<int, F> const& fl, int coarse)
            {
                double const* src = &m_cur[(start_from + fp) * FLAT];
                double        sum = 0.0;
                for (int i = 0; i < F; ++i) sum += src[fl[i]];
                m_cur[to * FLAT + coarse] = sum / static_cast<double>(F);
            }
        );
    }

    // ndtree::fragment(node) (ndtree.hpp:727-778)
    void fragment(key_t node)
    {
        const std::size_t from     = map_at(node);
        const std::size_t start_to = m_size;
        for (int i = 0; i < F; ++i)
        {
            const key_t child = m_morton.child_of(node, static_cast<unsigned>(i));
            if (m_map.count(child)) throw std::logic_error("child already exists");
            append(child, compute_child_neighbors(node, m_nbrs[from], i));
        }
        enforce_symmetry_after_refinement(node);
        interpolate_patch(from, start_to);
        m_map.erase(node);
    }

    // ndtree::recombine(parent) (ndtree.hpp:780-834)
    void recombine(key_t parent)
    {
        const key_t       c0    = m_morton.child_of(parent, 0);
        const std::size_t start = map_at(c0);
        std::array<nbrs_t, F> cn{};
        for (int i = 0; i < F; ++i)
        {
            const key_t       ci = m_morton.child_of(parent, static_cast<unsigned>(i));
            const std::size_t li = map_at(ci);
            if (li != start + static_cast<std::size_t>(i))
                throw std::logic_error("children are not contiguous");
            cn[i] = m_nbrs[li];
            m_map.erase(ci);
        }
        const nbrs_t      pn = compute_parent_neighbors(cn);
        const std::size_t to = m_size;
        append(parent, pn);
        restrict_patches(start, to);
        enforce_symmetry_after_recombining(parent, pn);
    }

    // ndtree::is_refine_elegible (ndtree.hpp:1887-1896)
    bool is_refine_eligible(std::size_t i) const
    {
        return m_status[i] == Refine && morton<R>::level(m_keys[i]) < m_morton.depth;
    }

    // ndtree::is_coarsen_elegible (ndtree.hpp:1898-1917)
    bool is_coarsen_eligible(key_t parent) const
    {
        for (int i = 0; i < F; ++i)
        {
            auto it = m_map.find(m_morton.child_of(parent, static_cast<unsigned>(i)));
            if (it == m_map.end()) return false;
            if (m_status[it->second] != Coarsen) return false;
        }
        return true;
    }

    // ndtree::apply_refine_coarsen (ndtree.hpp:904-940)
    void apply_refine_coarsen()
    {
        m_to_refine.clear();
        m_to_coarsen.clear();
        std::vector<key_t> parents;
        for (std::size_t i = 0; i < m_size; ++i)
        {
            if (m_keys[i] == 0) continue;
            parents.push_back(m_morton.parent_of(m_keys[i]));
        }
        std::sort(parents.begin(), parents.end());
        parents.erase(std::unique(parents.begin(), parents.end()), parents.end());
        for (std::size_t i = 0; i < m_size; ++i)
            if (is_refine_eligible(i)) m_to_refine.push_back(m_keys[i]);
        for (key_t p : parents)
        {
            if (!is_coarsen_eligible(p)) continue;
            // Recombining the root's children aborts in the reference
            // (compute_parent_neighbors, neighbor.hpp:505-540, yields finer
            // ids that enforce_symmetry_after_recombining looks up with
            // m_index_map.at after they were erased, ndtree.hpp:1106).
            if (p == 0) throw std::domain_error("coarsening to the root level is not supported");
            m_to_coarsen.push_back(p);
        }
    }

    // ndtree::balancing (ndtree.hpp:1127-1240)
    void balancing()
    {
        // refinement closure over a growing list (:1130-1165)
        for (std::size_t i = 0; i != m_to_refine.size(); ++i)
        {
            const key_t node = m_to_refine[i];
            for (int k = 0; k < NDIR; ++k)
            {
                const nbrs_t      na = m_nbrs[map_at(node)];
                neighbor_t const& n  = na[k];
                if (n.kind != rel::coarser) continue;
                if (std::find(m_to_refine.begin(), m_to_refine.end(), n.id) == m_to_refine.end())
                    m_to_refine.push_back(n.id);
            }
        }
        // level-descending order (:1167-1171)
        std::sort(
            m_to_refine.begin(), m_to_refine.end(),
            [](key_t a, key_t b) { return morton<R>::level(a) > morton<R>::level(b); }
        );
        // coarsen veto (:1173-1239)
        std::vector<key_t> remove;
        for (key_t parent : m_to_coarsen)
        {
            bool should_remove = false;
            for (int k = 0; k < NDIR && !should_remove; ++k)
            {
                const auto bc = boundary_children(k);
                for (int j = 0; j < K && !should_remove; ++j)
                {
                    const key_t       child = m_morton.child_of(parent, static_cast<unsigned>(bc[j]));
                    neighbor_t const& n     = m_nbrs[map_at(child)][k];
                    if (n.kind == rel::finer)
                        should_remove = true;
                    else if (n.kind == rel::same &&
                             std::find(m_to_refine.begin(), m_to_refine.end(), n.id) != m_to_refine.end())
                        should_remove = true;
                }
            }
            if (should_remove) remove.push_back(parent);
        }
        for (key_t id : remove)
            m_to_coarsen.erase(std::remove(m_to_coarsen.begin(), m_to_coarsen.end(), id), m_to_coarsen.end());
    }

    // ndtree::block_buffer_swap (ndtree.hpp:1979-2001)
    void block_buffer_swap(std::size_t i, std::size_t j)
    {
        if (i == j) return;
        std::swap(m_keys[i], m_keys[j]);
        std::swap(m_status[i], m_status[j]);
        std::swap(m_nbrs[i], m_nbrs[j]);
        std::swap_ranges(&m_cur[i * FLAT], &m_cur[i * FLAT] + FLAT, &m_cur[j * FLAT]);
    }

    // ndtree::compact (ndtree.hpp:1919-1948)
    void compact()
    {
        std::size_t tail = 0;
        for (std::size_t head = 0; head < m_size; ++head)
        {
            if (m_map.count(m_keys[head]))
            {
                block_buffer_swap(head, tail);
                ++tail;
            }
        }
        m_size = tail;
        m_map.clear();
        for (std::size_t i = 0; i < m_size; ++i) m_map[m_keys[i]] = i;
    }

    // ndtree::sort_buffers (ndtree.hpp:1335-1414): the in-place cycle walk is
    // a gather new[p] = old[order[p]] with order sorted by key.
    void sort_buffers()
    {
        std::vector<std::size_t> order(m_size);
        for (std::size_t i = 0; i < m_size; ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) { return m_keys[a] < m_keys[b]; });
        std::vector<key_t>       keys(m_size);
        std::vector<std::int8_t> status(m_size);
        std::vector<nbrs_t>      nbrs(m_size);
        std::vector<double>      data(m_size * FLAT);
        for (std::size_t p = 0; p < m_size; ++p)
        {
            const std::size_t s = order[p];
            keys[p]             = m_keys[s];
            status[p]           = m_status[s];
            nbrs[p]             = m_nbrs[s];
            std::copy_n(&m_cur[s * FLAT], FLAT, &data[p * FLAT]);
        }
        std::copy(keys.begin(), keys.end(), m_keys.begin());
        std::copy(status.begin(), status.end(), m_status.begin());
        std::copy(nbrs.begin(), nbrs.end(), m_nbrs.begin());
        std::copy(data.begin(), data.end(), m_cur.begin());
        m_map.clear();
        for (std::size_t i = 0; i < m_size; ++i) m_map[m_keys[i]] = i;
    }

    // ndtree::reconstruct_tree (ndtree.hpp:1249-1271) with per-leaf flags
    // given in linear order (update_refine_flags, ndtree.hpp:886-902).
    // Returns true when the topology changed.
    bool reconstruct_tree(std::int8_t const* flags)
    {
        for (std::size_t i = 0; i < m_size; ++i) m_status[i] = flags[i];
        apply_refine_coarsen();
        m_last_refine.clear();
        m_last_coarsen.clear();
        if (m_to_refine.empty() && m_to_coarsen.empty()) return false;
        balancing();
        m_last_refine  = m_to_refine;
        m_last_coarsen = m_to_coarsen;
        if (m_to_refine.empty() && m_to_coarsen.empty()) return false;
        for (std::size_t i = m_to_refine.size(); i > 0; --i) fragment(m_to_refine[i - 1]);
        for (key_t p : m_to_coarsen) recombine(p);
        compact();
        sort_buffers();
        return true;
    }

    // ndtree::neighbor_linear_index + rebuild_halo_exchange_metadata
    // (ndtree.hpp:691-725, 1606-1660)
    void export_metadata(halo_meta* out) const
    {
        for (std::size_t i = 0; i < m_size; ++i)
        {
            for (int k = 0; k < NDIR; ++k)
            {
                halo_meta         m{};
                neighbor_t const& n = m_nbrs[i][k];
                switch (n.kind)
                {
                    case rel::none: m.relation = 0; break;
                    case rel::same:
                        m.relation = 1;
                        m.neighbor = static_cast<std::int32_t>(map_at(n.id));
                        break;
                    case rel::finer:
                        m.relation = 2;
                        for (int j = 0; j < K; ++j) m.finer_ids[j] = static_cast<std::int32_t>(map_at(n.ids[j]));
                        break;
                    case rel::coarser:
                        m.relation = 3;
                        m.neighbor = static_cast<std::int32_t>(map_at(n.id));
                        for (int j = 0; j < R; ++j) m.quadrant[j] = n.quadrant[j];
                        break;
                }
                out[i * NDIR + k] = m;
            }
        }
    }

    // ---------------------------------------------------------------------
    // Halo exchange (patch_utils.hpp:63-447, ndtree.hpp:1581-1596)
    // ---------------------------------------------------------------------
    void halo_one(std::size_t i)
    {
        double* self = &m_cur[i * FLAT];
        for (int k = 0; k < NDIR; ++k)
        {
            const int  dim = dim_of(k);
            const bool pos = positive(k);
            // halo band of direction k, corners excluded (patch_layout.hpp:72-110)
            idx_t lo, hi;
            for (int d = 0; d < R; ++d)
            {
                lo[d] = (d != dim) ? H : (pos ? PS[d] - H : 0);
                hi[d] = (d != dim) ? PS[d] - H : (pos ? PS[d] : H);
            }
            neighbor_t const& n = m_nbrs[i][k];
            switch (n.kind)
            {
                case rel::none: break; // boundary_t: no-op (patch_utils.hpp:303-313)
                case rel::same:
                {
                    // same_t (patch_utils.hpp:315-332)
                    double const* other = &m_cur[map_at(n.id) * FLAT];
                    box_for(
                        lo, hi,
                        [&](idx_t const& idxs)
                        {
                            idx_t from = idxs;
                            from[dim]  = pos ? from[dim] - NS[dim] : from[dim] + NS[dim];
                            self[lin(idxs)] = other[lin(from)];
                        }
                    );
                    break;
                }
                case rel::finer:
                {
                    // finer_t (patch_utils.hpp:334-386)
                    std::array<double const*, K> fines{};
                    for (int j = 0; j < K; ++j) fines[j] = &m_cur[map_at(n.ids[j]) * FLAT];
                    box_for(
                        lo, hi,
                        [&](idx_t const& idxs)
                        {
                            int   fine_patch = 0, stride = 1;
                            idx_t base = idxs;
                            base[dim]  = pos ? base[dim] - NS[dim] : base[dim] + NS[dim];
                            for (int d = 0; d < R; ++d)
                            {
                                // each dim reduced modulo its own extent, as
                                // the device twin does
                                // (src/cuda/halo_exchange.cu:174-176); the
                                // CPU rule takes the normal dim's extent for
                                // every d (:358), the same for square
                                // patches only
                                base[d] = ((base[d] - H) * 2) % NS[d] + H;
                                if (d != dim)
                                {
                                    const int section = NS[d] / 2;
                                    fine_patch += ((idxs[d] - H) / section) * stride;
                                    stride *= 2;
                                }
                            }
                            double const* fp  = fines[fine_patch];
                            double        sum = 0.0;
                            idx_t         zero{}, two;
                            two.fill(2);
                            box_for(
                                zero, two,
                                [&](idx_t const& off)
                                {
                                    idx_t c = base;
                                    for (int d = 0; d < R; ++d) c[d] += off[d];
                                    sum += fp[lin(c)];
                                }
                            );
                            self[lin(idxs)] = sum / static_cast<double>(F);
                        }
                    );
                    break;
                }
                case rel::coarser:
                {
                    // coarser_t (patch_utils.hpp:388-441): injection
                    double const* other = &m_cur[map_at(n.id) * FLAT];
                    box_for(
                        lo, hi,
                        [&](idx_t const& idxs)
                        {
                            idx_t from = idxs;
                            from[dim]  = pos ? from[dim] - NS[dim] : from[dim] + NS[dim];
                            for (int d = 0; d < R; ++d)
                            {
                                const int fine_mapped = from[d] - H;
                                from[d] = H + n.quadrant[d] * (NS[d] / 2) + fine_mapped / 2;
                            }
                            self[lin(idxs)] = other[lin(from)];
                        }
                    );
                    break;
                }
            }
        }
    }

    void halo_exchange_update()
    {
        parallel_for(m_size, m_threads, [&](std::size_t i) { halo_one(i); });
    }

    // ---------------------------------------------------------------------
    // Solver (amr_solver.hpp:149-413, AdvectionPhysics.hpp:45-85) with the
    // velocity as a run-time parameter.
    // ---------------------------------------------------------------------
    // rusanovFlux (AdvectionPhysics.hpp:45-66)
    double flux(double ul, double ur, int d) const { return advection_flux(ul, ur, m_vel[d]); }

    // compute_time_step_cpu (amr_solver.hpp:355-413)
    double compute_time_step() const
    {
        double dt = std::numeric_limits<double>::max();
        for (std::size_t p = 0; p < m_size; ++p)
        {
            const auto cs = cell_sizes(m_keys[p]);
            for (int d = 0; d < R; ++d)
            {
                const double speed = std::abs(m_vel[d]);
                if (speed > 1e-12) dt = std::min(dt, cs[d] / speed);
            }
        }
        return m_cfl * dt;
    }

    // time_step_cpu (amr_solver.hpp:265-353) without the trailing swap /
    // halo exchange.  Physical dim d uses stride of layout dim R-1-d.
    void time_step_kernel(double dt)
    {
        parallel_for(
            m_size, m_threads,
            [&](std::size_t p)
            {
                const auto    cs  = cell_sizes(m_keys[p]);
                double const* in  = &m_cur[p * FLAT];
                double*       out = &m_next[p * FLAT];
                interior_for(
                    [&](idx_t const& idxs)
                    {
                        const int    l     = lin(idxs);
                        double       total = 0.0;
                        const double uc    = in[l];
                        for (int d = 0; d < R; ++d)
                        {
                            const int    s    = S[R - 1 - d];
                            const double ul   = in[l - s];
                            const double ur   = in[l + s];
                            const double fL   = flux(ul, uc, d);
                            const double fR   = flux(uc, ur, d);
                            const double dtdx = dt / cs[d];
                            total -= dtdx * (fR - fL);
                        }
                        out[l] = in[l] + total;
                    }
                );
            }
        );
    }

    void swap_buffers() { std::swap(m_cur, m_next); }

    void time_step(double dt)
    {
        time_step_kernel(dt);
        swap_buffers();
        halo_exchange_update();
    }

    // advance_batch_async / finish_advance_batch CPU branch
    // (amr_solver.hpp:221-262)
    double advance_batch(std::size_t steps, double remaining, std::size_t* executed)
    {
        double      acc = 0.0;
        std::size_t cnt = 0;
        for (std::size_t s = 0; s < steps; ++s)
        {
            if (remaining <= 0.0) break;
            const double dt      = compute_time_step();
            const double step_dt = std::min(dt, remaining);
            if (step_dt <= 0.0) break;
            time_step(step_dt);
            acc += step_dt;
            remaining -= step_dt;
            ++cnt;
        }
        if (executed) *executed = cnt;
        return acc;
    }

    // ---------------------------------------------------------------------
    // Refinement criteria producing per-leaf flags.
    // ---------------------------------------------------------------------
    // Max-value criterion of the reference's examples/tests
    // (examples/fvm_solver_advection.e.cpp:66-101,
    // tests/advection_equation_amr.t.cpp:155-170): max starts at 0 and runs
    // over the padded patch; corner cells are excluded here because no
    // halo exchange ever writes them.
    void criterion_max_value(double refine_thr, double coarsen_thr, int min_level, int max_level,
                             std::int8_t* flags) const
    {
        for (std::size_t p = 0; p < m_size; ++p)
        {
            double const* u  = &m_cur[p * FLAT];
            double        mx = 0.0;
            idx_t         lo{}, hi = PS;
            box_for(
                lo, hi,
                [&](idx_t const& idxs)
                {
                    int outside = 0;
                    for (int d = 0; d < R; ++d) outside += (idxs[d] < H || idxs[d] >= H + NS[d]) ? 1 : 0;
                    if (outside <= 1) mx = std::max(mx, u[lin(idxs)]);
                }
            );
            const int l = morton<R>::level(m_keys[p]);
            if (mx > refine_thr && l < max_level)
                flags[p] = Refine;
            else if (mx < coarsen_thr && l > min_level)
                flags[p] = Coarsen;
            else
                flags[p] = Stable;
        }
    }

    // Undivided-gradient criterion (the reference has none; defined here):
    // g = max over interior cells c and layout dims of
    // max(|u[c+s]-u[c]|, |u[c]-u[c-s]|) using the exchanged face halos.
    void criterion_gradient(double refine_thr, double coarsen_thr, int min_level, int max_level,
                            std::int8_t* flags) const
    {
        for (std::size_t p = 0; p < m_size; ++p)
        {
            double const* u = &m_cur[p * FLAT];
            double        g = 0.0;
            interior_for(
                [&](idx_t const& idxs)
                {
                    const int l = lin(idxs);
                    for (int d = 0; d < R; ++d)
                    {
                        const double a = std::abs(u[l + S[d]] - u[l]);
                        const double b = std::abs(u[l] - u[l - S[d]]);
                        g              = std::max(g, std::max(a, b));
                    }
                }
            );
            const int lv = morton<R>::level(m_keys[p]);
            if (g > refine_thr && lv < max_level)
                flags[p] = Refine;
            else if (g < coarsen_thr && lv > min_level)
                flags[p] = Coarsen;
            else
                flags[p] = Stable;
        }
    }

private:
    morton<R>                             m_morton;
    std::size_t                           m_capacity;
    std::array<double, 3>                 m_lengths;
    std::size_t                           m_size = 0;
    std::vector<key_t>                    m_keys;
    std::vector<std::int8_t>              m_status;
    std::vector<nbrs_t>                   m_nbrs;
    std::unordered_map<key_t, std::size_t> m_map;
    std::vector<double>                   m_cur, m_next;
    std::vector<key_t>                    m_to_refine, m_to_coarsen;
    std::vector<key_t>                    m_last_refine, m_last_coarsen;
    int                                   m_threads = 1;
    std::array<double, 3>                 m_vel{ 1.0, 0.5, 0.0 };
    double                                m_cfl = 0.4;
};

// ---------------------------------------------------------------------------
// Euler physics (include/solver/EulerPhysics.hpp) and the amr_solver step /
// dt over its NV = R + 2 conservative fields.  The reference keeps one SoA
// patch per field in one tree (FieldTags, EulerPhysics.hpp:22-30); here each
// field lives in its own tree and all trees go through the same regrids, so
// their topologies stay identical.
// ---------------------------------------------------------------------------
template <int R>
struct euler_physics
{
    static constexpr int NV = R + 2;

    // primitiveToConservative (EulerPhysics.hpp:39-62): [rho, u.., p] ->
    // [rho, rho u.., E], E = p / (gamma - 1) + 0.5 rho |u|^2.
    static void prim_to_cons(double const* prim, double* cons, double gamma)
    {
        const double rho = prim[0];
        const double p   = prim[R + 1];
        cons[0]          = rho;
        double vel_sq    = 0.0;
        for (int d = 0; d < R; ++d)
        {
            const double u = prim[1 + d];
            cons[1 + d]    = rho * u;
            vel_sq += u * u;
        }
        cons[R + 1] = p / (gamma - 1.0) + 0.5 * rho * vel_sq;
    }

    // The per-state terms of rusanovFlux / getMaxSpeed
    // (EulerPhysics.hpp:80-94, 140-152).
    struct terms
    {
        double p, a, u_dir;
    };
    static terms state(double const* U, int dir, double gamma)
    {
        const double inv_rho = 1.0 / U[0];
        double       K       = 0.0;
        for (int d = 0; d < R; ++d) K += U[1 + d] * U[1 + d];
        K *= 0.5 * inv_rho;
        const double p = (gamma - 1.0) * (U[R + 1] - K);
        const double a = std::sqrt(gamma * p * inv_rho);
        return { p, a, U[1 + dir] * inv_rho };
    }

    // std::max as the reference calls it: (a < b) ? b : a.
    static double max_of(double a, double b) { return (a < b) ? b : a; }

    // rusanovFlux (EulerPhysics.hpp:72-125).
    static void rusanov(double const* UL, double const* UR, double* flux, int dir, double gamma)
    {
        const terms  L    = state(UL, dir, gamma);
        const terms  Rr   = state(UR, dir, gamma);
        const double smax = max_of(std::abs(L.u_dir) + L.a, std::abs(Rr.u_dir) + Rr.a);
        flux[0]           = 0.5 * (UL[1 + dir] + UR[1 + dir] - smax * (UR[0] - UL[0]));
        for (int d = 0; d < R; ++d)
        {
            double fl = UL[1 + d] * L.u_dir;
            double fr = UR[1 + d] * Rr.u_dir;
            if (d == dir)
            {
                fl += L.p;
                fr += Rr.p;
            }
            flux[1 + d] = 0.5 * (fl + fr - smax * (UR[1 + d] - UL[1 + d]));
        }
        const double fle = L.u_dir * (UL[R + 1] + L.p);
        const double fre = Rr.u_dir * (UR[R + 1] + Rr.p);
        flux[R + 1]      = 0.5 * (fle + fre - smax * (UR[R + 1] - UL[R + 1]));
    }

    // getMaxSpeed (EulerPhysics.hpp:132-155).
    static double max_speed(double const* U, int dir, double gamma)
    {
        const terms t = state(U, dir, gamma);
        return std::abs(t.u_dir) + t.a;
    }
};

template <int R, int H, int... Ns>
class euler_system
{
  public:
    using tree_t               = tree<R, H, Ns...>;
    using phys                 = euler_physics<R>;
    static constexpr int NV    = phys::NV;
    static constexpr int FLAT  = tree_t::FLAT;
    static constexpr int NI    = tree_t::NI;

    euler_system(int depth, std::size_t capacity, std::array<double, 3> lengths)
    {
        for (auto& f : m_f) f = std::make_unique<tree_t>(depth, capacity, lengths);
    }

    tree_t&       field(int v) { return *m_f[v]; }
    tree_t const& field(int v) const { return *m_f[v]; }
    std::size_t   size() const { return m_f[0]->size(); }
    void          set_params(double gamma, double cfl, int threads)
    {
        m_gamma = gamma;
        m_cfl   = cfl;
        m_threads = std::max(1, threads);
        for (auto& f : m_f) f->set_threads(m_threads);
    }

    // ndtree::reconstruct_tree on every field (the flags fix the topology).
    bool reconstruct(std::int8_t const* flags)
    {
        bool changed = false;
        for (auto& f : m_f) changed = f->reconstruct_tree(flags);
        return changed;
    }
    void halo_exchange_update()
    {
        for (auto& f : m_f) f->halo_exchange_update();
    }

    // compute_time_step_cpu (amr_solver.hpp:355-413) with getMaxSpeed.
    double compute_time_step() const
    {
        std::vector<double> local(size(), std::numeric_limits<double>::max());
        parallel_for(
            size(), m_threads,
            [&](std::size_t p)
            {
                const auto cs = m_f[0]->cell_sizes(m_f[0]->key_at(p));
                double     dt = std::numeric_limits<double>::max();
                tree_t::interior_for(
                    [&](typename tree_t::idx_t const& idx)
                    {
                        const int l = tree_t::lin(idx);
                        double    U[NV];
                        for (int v = 0; v < NV; ++v) U[v] = m_f[v]->patch(p)[l];
                        for (int d = 0; d < R; ++d)
                        {
                            const double speed = phys::max_speed(U, d, m_gamma);
                            if (speed > 1e-12) dt = std::min(dt, cs[d] / speed);
                        }
                    }
                );
                local[p] = dt;
            }
        );
        double dt = std::numeric_limits<double>::max();
        for (double x : local) dt = std::min(dt, x);
        return m_cfl * dt;
    }

    // time_step_cpu (amr_solver.hpp:265-353): per cell and physical dim d,
    // fL = F(U[c - s_d], U[c]), fR = F(U[c], U[c + s_d]);
    // total[k] -= dt / dx_d * (fR[k] - fL[k]); out = U + total.
    void time_step_kernel(double dt)
    {
        parallel_for(
            size(), m_threads,
            [&](std::size_t p)
            {
                const auto    cs = m_f[0]->cell_sizes(m_f[0]->key_at(p));
                double const* in[NV];
                double*       out[NV];
                for (int v = 0; v < NV; ++v)
                {
                    in[v]  = m_f[v]->patch(p);
                    out[v] = m_f[v]->next_patch(p);
                }
                tree_t::interior_for(
                    [&](typename tree_t::idx_t const& idx)
                    {
                        const int l = tree_t::lin(idx);
                        double    total[NV] = {};
                        double    Uc[NV];
                        for (int v = 0; v < NV; ++v) Uc[v] = in[v][l];
                        for (int d = 0; d < R; ++d)
                        {
                            const int s = tree_t::S[R - 1 - d];
                            double    UL[NV], UR[NV], fL[NV], fR[NV];
                            for (int v = 0; v < NV; ++v)
                            {
                                UL[v] = in[v][l - s];
                                UR[v] = in[v][l + s];
                            }
                            phys::rusanov(UL, Uc, fL, d, m_gamma);
                            phys::rusanov(Uc, UR, fR, d, m_gamma);
                            const double dt_over_dx = dt / cs[d];
                            for (int k = 0; k < NV; ++k) total[k] -= dt_over_dx * (fR[k] - fL[k]);
                        }
                        for (int v = 0; v < NV; ++v) out[v][l] = in[v][l] + total[v];
                    }
                );
            }
        );
    }

    void time_step(double dt)
    {
        time_step_kernel(dt);
        for (auto& f : m_f) f->swap_buffers();
        halo_exchange_update();
    }

    // advance_batch_async / finish_advance_batch CPU branch
    // (amr_solver.hpp:221-262).
    double advance_batch(std::size_t steps, double remaining, std::size_t* executed)
    {
        double      acc = 0.0;
        std::size_t cnt = 0;
        for (std::size_t s = 0; s < steps; ++s)
        {
            if (remaining <= 0.0) break;
            const double dt      = compute_time_step();
            const double step_dt = std::min(dt, remaining);
            if (step_dt <= 0.0) break;
            time_step(step_dt);
            acc += step_dt;
            remaining -= step_dt;
            ++cnt;
        }
        if (executed) *executed = cnt;
        return acc;
    }

  private:
    std::array<std::unique_ptr<tree_t>, NV> m_f;
    double                                  m_gamma   = 1.4;
    double                                  m_cfl     = 0.3;
    int                                     m_threads = 1;
};

} // namespace oracle
