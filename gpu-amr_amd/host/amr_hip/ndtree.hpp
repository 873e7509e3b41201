// ndtree.hpp -- C++ host API of the MI355X patch-AMR advection step.
//
// Mirrors the names and roles of the reference's template API so that code
// written against it ports by changing types only:
//   amr::ndt::morton::morton_id<Depth, Rank>   (include/morton/morton_id.hpp:20-450)
//   amr::ndt::tree::ndtree<...>                (include/ndtree/ndtree.hpp:94-2157)
//   amr_solver<Tree, Geometry, Equation, DIM>  (include/solver/amr_solver.hpp:28-544)
//     -- advection: amr_solver<Tree>; Euler: euler_solver<Tree>
//   vtk_print<Physics_System>                  (include/ndtree/vtk_print.hpp:16-280)
// Every call goes through the C ABI of include/amr_hip.h; errors surface as
// std::runtime_error, as the reference's device layer reports them
// (src/cuda/device_buffer.cu:17-27).
#pragma once

#include "amr_hip.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

namespace amr::hip
{

inline void check(int status, char const* what)
{
    if (status != AMR_OK) throw std::runtime_error(std::string(what) + ": " + amr_last_error());
}

// Morton patch index: key = (interleave(x, y[, z]) << 6) | level.
template <unsigned Depth, unsigned Rank>
class morton_id
{
    static_assert(Rank == 2 || Rank == 3);

public:
    using mask_t      = std::uint64_t;
    using coord_array = std::array<std::uint32_t, Rank>;
    using offset_t    = std::uint32_t;
    using level_t     = std::uint8_t;

    constexpr morton_id() = default;
    constexpr morton_id(mask_t id) : m_id{ id } {}

    constexpr mask_t id() const noexcept { return m_id; }
    constexpr level_t level() const noexcept { return static_cast<level_t>(m_id & 0x3F); }
    static constexpr level_t level(morton_id i) { return i.level(); }
    static constexpr unsigned max_depth() { return Depth; }
    static constexpr unsigned rank() { return Rank; }
    static constexpr morton_id root() { return morton_id(0); }

    static constexpr mask_t encode(coord_array const& c, level_t level)
    {
        mask_t m = 0;
        for (unsigned b = 0; b < Depth; ++b)
            for (unsigned i = 0; i < Rank; ++i) m |= static_cast<mask_t>((c[i] >> b) & 1u) << (Rank * b + i);
        return (m << 6) | (level & 0x3F);
    }
    static constexpr std::pair<coord_array, level_t> decode(mask_t id)
    {
        coord_array c{};
        const mask_t m = id >> 6;
        for (unsigned b = 0; b < Depth; ++b)
            for (unsigned i = 0; i < Rank; ++i) c[i] |= static_cast<std::uint32_t>((m >> (Rank * b + i)) & 1u) << b;
        return { c, static_cast<level_t>(id & 0x3F) };
    }
    static constexpr morton_id parent_of(morton_id k)
    {
        auto [c, l]            = decode(k.id());
        const std::uint32_t s  = 1u << (Depth - l);
        for (auto& x : c) x &= ~s;
        return morton_id(encode(c, static_cast<level_t>(l - 1)));
    }
    static constexpr morton_id child_of(morton_id p, offset_t o)
    {
        auto [c, l]           = decode(p.id());
        const std::uint32_t s = 1u << (Depth - l - 1);
        for (unsigned i = 0; i < Rank; ++i) c[i] += ((o >> i) & 1u) * s;
        return morton_id(encode(c, static_cast<level_t>(l + 1)));
    }
    static constexpr offset_t offset_of(morton_id k)
    {
        auto [c, l] = decode(k.id());
        if (l == 0) return 0;
        offset_t o = 0;
        for (unsigned i = 0; i < Rank; ++i) o |= ((c[i] >> (Depth - l)) & 1u) << i;
        return o;
    }

    constexpr bool operator==(morton_id const&) const noexcept = default;
    constexpr bool operator<(morton_id const& o) const noexcept { return m_id < o.m_id; }

private:
    mask_t m_id = 0;
};

enum class refine_status_t : std::int8_t
{
    Stable  = 0,
    Refine  = 1,
    Coarsen = 2,
};

// Device-evaluated criterion (the fill_refine_flags form, ndtree.hpp:889-894).
struct device_criterion
{
    amr_criterion c{};
    static device_criterion max_value(double refine, double coarsen, int min_level, int max_level)
    {
        return { { AMR_CRITERION_MAX_VALUE, min_level, max_level, 0, refine, coarsen } };
    }
    static device_criterion gradient(double refine, double coarsen, int min_level, int max_level)
    {
        return { { AMR_CRITERION_GRADIENT, min_level, max_level, 0, refine, coarsen } };
    }
};

// static_shape<N, M[, O]> (include/containers/static_shape.hpp,
// include/containers/static_layout.hpp:28-37): interior cells per layout
// dim, row-major, the last dim is physical x.
template <unsigned... Ns>
struct static_shape
{
    static constexpr unsigned                            rank = sizeof...(Ns);
    static constexpr std::array<unsigned, sizeof...(Ns)> sizes{ Ns... };
    static constexpr std::size_t                         elements = (std::size_t{ 1 } * ... * Ns);
    static_assert(rank == 2 || rank == 3, "2D or 3D patches");
};

namespace detail
{
template <unsigned Rank, unsigned N>
struct cube;
template <unsigned N>
struct cube<2, N>
{
    using type = static_shape<N, N>;
};
template <unsigned N>
struct cube<3, N>
{
    using type = static_shape<N, N, N>;
};
} // namespace detail

// basic_ndtree<Depth, Shape, Halo>: a periodic tree of leaf patches of
// Shape (a static_shape) with Halo ghost layers in the padded view.
template <unsigned Depth, typename Shape, unsigned Halo>
class basic_ndtree
{
public:
    static constexpr unsigned Rank = Shape::rank;
    using shape_t         = Shape;
    using patch_index_t   = morton_id<Depth, Rank>;
    using linear_index_t  = std::size_t;
    using refine_status_t = amr::hip::refine_status_t;
    static constexpr unsigned    s_rank     = Rank;
    static constexpr std::size_t s_interior = Shape::elements;
    // padded extent per layout dim
    static constexpr std::array<unsigned, Rank> s_padded = []
    {
        std::array<unsigned, Rank> p{};
        for (unsigned d = 0; d < Rank; ++d) p[d] = Shape::sizes[d] + 2 * Halo;
        return p;
    }();
    static constexpr std::size_t s_flat = []
    {
        std::size_t f = 1;
        for (unsigned d : Shape::sizes) f *= d + 2 * Halo;
        return f;
    }();

    explicit basic_ndtree(std::size_t capacity, std::array<double, 3> lengths = { 1.0, 1.0, 1.0 })
    {
        const auto cfg = config(capacity, lengths);
        check(amr_tree_create(&cfg, &m_ctx), "ndtree");
        m_lengths = lengths;
        refresh_keys();
    }
    // The tree of an Euler solver: Rank + 2 conservative fields per cell
    // (EulerPhysics<Rank>::FieldTags, EulerPhysics.hpp:22-30).
    struct euler_fields
    {
        double gamma, cfl;
    };
    basic_ndtree(std::size_t capacity, std::array<double, 3> lengths, euler_fields e)
    {
        const auto cfg = config(capacity, lengths);
        check(amr_euler_create(&cfg, e.gamma, e.cfl, &m_ctx), "ndtree (Euler)");
        m_lengths = lengths;
        m_fields  = Rank + 2;
        refresh_keys();
    }
    ~basic_ndtree()
    {
        if (m_ctx) amr_tree_destroy(m_ctx);
    }
    basic_ndtree(basic_ndtree const&)            = delete;
    basic_ndtree& operator=(basic_ndtree const&) = delete;

    amr_ctx* handle() const noexcept { return m_ctx; }
    unsigned fields() const noexcept { return m_fields; }

    std::size_t size() const noexcept { return m_keys.size(); }
    patch_index_t get_node_index_at(linear_index_t i) const { return patch_index_t(m_keys.at(i)); }

    // reconstruct_tree(fn): fn(patch_index_t) -> refine_status_t evaluated
    // per leaf on the host, or a device_criterion evaluated on the device.
    template <typename Fn>
    bool reconstruct_tree(Fn&& fn)
    {
        int changed = 0;
        if constexpr (std::is_same_v<std::decay_t<Fn>, device_criterion>)
            check(amr_tree_reconstruct_criterion(m_ctx, &fn.c, &changed), "reconstruct_tree");
        else
        {
            std::vector<std::int8_t> flags(size());
            for (std::size_t i = 0; i < size(); ++i)
                flags[i] = static_cast<std::int8_t>(fn(get_node_index_at(i)));
            check(amr_tree_reconstruct_flags(m_ctx, flags.data(), flags.size(), &changed), "reconstruct_tree");
        }
        refresh_keys();
        m_view_valid = false;
        return changed != 0;
    }

    void halo_exchange_update()
    {
        check(amr_tree_halo_exchange_update(m_ctx), "halo_exchange_update");
        m_view_valid = false;
    }
    void swap_buffers()
    {
        check(amr_tree_swap_buffers(m_ctx), "swap_buffers");
        m_view_valid = false;
    }

    // Padded patch (prod (n_d + 2 Halo) doubles, halos exchanged, corners 0) of
    // field `field` (get_patch<Field>, field index in FieldTags order) by
    // linear index or by patch index (ndtree.hpp:477-512).  The padded view
    // is downloaded once and reused until the state changes.
    std::vector<double> get_patch(linear_index_t i, unsigned field = 0) const
    {
        auto const&       all = view();
        const std::size_t at  = (static_cast<std::size_t>(field) * size() + i) * s_flat;
        return std::vector<double>(all.begin() + static_cast<std::ptrdiff_t>(at),
                                   all.begin() + static_cast<std::ptrdiff_t>(at + s_flat));
    }
    std::vector<double> get_patch(patch_index_t k, unsigned field = 0) const
    {
        return get_patch(linear_index_of(k), field);
    }
    std::vector<double> get_patches() const { return view(); }
    linear_index_t linear_index_of(patch_index_t k) const
    {
        auto it = std::lower_bound(m_keys.begin(), m_keys.end(), k.id());
        if (it == m_keys.end() || *it != k.id()) throw std::out_of_range("patch index is not a leaf");
        return static_cast<linear_index_t>(it - m_keys.begin());
    }
    void invalidate_view() noexcept { m_view_valid = false; }
    // All fields' interiors, field-major (field, leaf, cell).
    std::vector<double> get_interior() const
    {
        std::vector<double> out(m_fields * size() * s_interior);
        check(amr_tree_get_interior(m_ctx, out.data(), out.size()), "get_interior");
        return out;
    }
    void set_interior(std::vector<double> const& v)
    {
        check(amr_tree_set_interior(m_ctx, v.data(), v.size()), "set_interior");
        m_view_valid = false;
    }
    std::vector<amr_halo_direction_metadata> neighbor_metadata() const
    {
        std::vector<amr_halo_direction_metadata> out(size() * 2 * Rank);
        check(amr_tree_get_neighbor_metadata(m_ctx, out.data(), out.size()), "neighbor_metadata");
        return out;
    }

    // physics_system (include/solver/physics_system.hpp:150-231)
    std::array<double, Rank> cell_sizes(patch_index_t k) const
    {
        std::array<double, Rank> r{};
        const double             pm   = static_cast<double>(1u << (Depth - k.level()));
        const double             maxc = static_cast<double>(1u << Depth);
        for (unsigned i = 0; i < Rank; ++i)
            r[i] = (m_lengths[i] * pm / maxc) / static_cast<double>(Shape::sizes[Rank - 1 - i]);
        return r;
    }
    // Centre of interior cell (row-major interior index ci) of leaf k.
    std::array<double, Rank> cell_center(patch_index_t k, std::size_t ci) const
    {
        auto [c, l]    = patch_index_t::decode(k.id());
        const auto cs  = cell_sizes(k);
        const double maxc = static_cast<double>(1u << Depth);
        std::array<double, Rank> r{};
        std::size_t              rem = ci;
        std::array<std::size_t, Rank> idx{};
        for (int d = static_cast<int>(Rank) - 1; d >= 0; --d)
        {
            idx[static_cast<std::size_t>(d)] = rem % Shape::sizes[static_cast<std::size_t>(d)];
            rem /= Shape::sizes[static_cast<std::size_t>(d)];
        }
        for (unsigned i = 0; i < Rank; ++i)
        {
            const double origin = m_lengths[i] * static_cast<double>(c[i]) / maxc;
            const double cell   = origin + static_cast<double>(idx[Rank - 1 - i]) * cs[i];
            r[i]                = cell + 0.5 * cs[i];
        }
        return r;
    }

private:
    static amr_tree_config config(std::size_t capacity, std::array<double, 3> const& lengths)
    {
        amr_tree_config cfg{};
        cfg.rank       = static_cast<int32_t>(Rank);
        cfg.max_depth  = static_cast<int32_t>(Depth);
        for (unsigned d = 0; d < Rank; ++d) cfg.patch_extents[d] = static_cast<int32_t>(Shape::sizes[d]);
        cfg.halo_width = static_cast<int32_t>(Halo);
        cfg.capacity   = capacity;
        for (int i = 0; i < 3; ++i) cfg.domain_length[i] = lengths[i];
        return cfg;
    }

    std::vector<double> const& view() const
    {
        if (!m_view_valid)
        {
            m_view.resize(m_fields * size() * s_flat);
            check(amr_tree_get_patches(m_ctx, m_view.data(), m_view.size()), "get_patches");
            m_view_valid = true;
        }
        return m_view;
    }

    void refresh_keys()
    {
        std::uint64_t n = 0;
        check(amr_tree_size(m_ctx, &n), "size");
        m_keys.resize(n);
        check(amr_tree_get_keys(m_ctx, m_keys.data(), n), "get_keys");
    }

    amr_ctx*                    m_ctx    = nullptr;
    unsigned                    m_fields = 1;
    std::vector<std::uint64_t>  m_keys;
    std::array<double, 3>       m_lengths{ 1.0, 1.0, 1.0 };
    mutable std::vector<double> m_view;
    mutable bool                m_view_valid = false;
};

// ndtree<Rank, Depth, N, Halo>: square / cubic N^Rank-cell patches.
template <unsigned Rank, unsigned Depth, unsigned N, unsigned Halo>
using ndtree = basic_ndtree<Depth, typename detail::cube<Rank, N>::type, Halo>;

// amr_solver<Tree> for linear advection with a run-time velocity.
template <typename Tree>
class amr_solver
{
public:
    amr_solver(std::size_t capacity, [[maybe_unused]] double gamma, double cfl,
               std::array<double, 3> velocity = { 1.0, 0.5, 0.0 }, std::array<double, 3> lengths = { 1.0, 1.0, 1.0 })
        : m_tree(capacity, lengths)
    {
        check(amr_solver_configure(m_tree.handle(), velocity.data(), cfl), "amr_solver");
    }

    Tree& get_tree() { return m_tree; }

    // initialize(ic): ic(std::array<double, Rank> centre) -> double
    template <typename IC>
    void initialize(IC&& ic)
    {
        std::vector<double> v(m_tree.size() * Tree::s_interior);
        for (std::size_t p = 0; p < m_tree.size(); ++p)
        {
            const auto k = m_tree.get_node_index_at(p);
            for (std::size_t c = 0; c < Tree::s_interior; ++c) v[p * Tree::s_interior + c] = ic(m_tree.cell_center(k, c));
        }
        m_tree.set_interior(v);
    }

    double advance()
    {
        double dt = 0.0;
        check(amr_solver_advance(m_tree.handle(), &dt), "advance");
        m_tree.invalidate_view();
        return dt;
    }
    void advance_batch_async(std::size_t steps, double remaining = std::numeric_limits<double>::max())
    {
        check(amr_solver_advance_batch_async(m_tree.handle(), steps, remaining), "advance_batch_async");
        m_tree.invalidate_view();
    }
    double finish_advance_batch(std::size_t* executed = nullptr)
    {
        double        dt = 0.0;
        std::uint64_t n  = 0;
        check(amr_solver_finish_advance_batch(m_tree.handle(), &dt, &n), "finish_advance_batch");
        if (executed) *executed = static_cast<std::size_t>(n);
        return dt;
    }

private:
    Tree m_tree;
};

// amr_solver<Tree, physics_system, EulerPhysics<Rank>, Rank>(capacity, gamma,
// cfl) (amr_solver.hpp:28-97, include/solver/EulerPhysics.hpp): the tree
// carries [rho, rho u, rho v, (rho w), E]; dt follows the maximum wave speed
// every step, on the device.
template <typename Tree>
class euler_solver
{
public:
    static constexpr unsigned s_nvar = Tree::s_rank + 2;

    euler_solver(std::size_t capacity, double gamma, double cfl, std::array<double, 3> lengths = { 1.0, 1.0, 1.0 })
        : m_tree(capacity, lengths, typename Tree::euler_fields{ gamma, cfl })
    {
    }

    Tree& get_tree() { return m_tree; }

    // initialize(ic): ic(std::array<double, Rank> centre) -> primitive
    // std::array<double, Rank + 2> [rho, u, v, (w), p]; converted by
    // primitiveToConservative (amr_solver.hpp:105-145).
    template <typename IC>
    void initialize(IC&& ic)
    {
        const std::size_t   per = m_tree.size() * Tree::s_interior;
        std::vector<double> prim(per * s_nvar);
        for (std::size_t p = 0; p < m_tree.size(); ++p)
        {
            const auto k = m_tree.get_node_index_at(p);
            for (std::size_t c = 0; c < Tree::s_interior; ++c)
            {
                const auto v = ic(m_tree.cell_center(k, c));
                for (unsigned f = 0; f < s_nvar; ++f) prim[f * per + p * Tree::s_interior + c] = v[f];
            }
        }
        check(amr_euler_set_primitive(m_tree.handle(), prim.data(), prim.size()), "initialize");
        m_tree.invalidate_view();
    }

    double compute_time_step()
    {
        double dt = 0.0;
        check(amr_solver_compute_dt(m_tree.handle(), &dt), "compute_time_step");
        return dt;
    }
    double advance()
    {
        double dt = 0.0;
        check(amr_solver_advance(m_tree.handle(), &dt), "advance");
        m_tree.invalidate_view();
        return dt;
    }
    void advance_batch_async(std::size_t steps, double remaining = std::numeric_limits<double>::max())
    {
        check(amr_solver_advance_batch_async(m_tree.handle(), steps, remaining), "advance_batch_async");
        m_tree.invalidate_view();
    }
    double finish_advance_batch(std::size_t* executed = nullptr)
    {
        double        dt = 0.0;
        std::uint64_t n  = 0;
        check(amr_solver_finish_advance_batch(m_tree.handle(), &dt, &n), "finish_advance_batch");
        if (executed) *executed = static_cast<std::size_t>(n);
        return dt;
    }

private:
    Tree m_tree;
};

// vtk_print<physics_system> (include/ndtree/vtk_print.hpp:16-62): files go
// to vtk_output/<base><extension>, the directory created on construction.
class vtk_print
{
public:
    explicit vtk_print(std::string_view base_filename)
        : m_base(base_filename)
    {
        std::filesystem::create_directory("vtk_output");
    }
    template <typename Tree>
    void print(Tree const& tree, std::string const& extension) const
    {
        const std::string path = "vtk_output/" + m_base + extension;
        check(amr_tree_write_vtk(tree.handle(), path.c_str()), "vtk_print");
    }

private:
    std::string m_base;
};

} // namespace amr::hip
