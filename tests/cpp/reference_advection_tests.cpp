// The reference's advection tests (tests/advection_equation_amr.t.cpp:50-231)
// written against the amr::hip C++ host API: same configuration (10^2 and
// 50^2 patches, halo 2, morton_id<7, 2>, CFL 0.4, pulse exp(-r^2/0.005) from
// (0.2, 0.2) to t = 0.5), same checks (argmax within one cell of x0 + v t).
// gtest is not available offline, so each case prints PASS/FAIL and the exit
// status counts failures.
#include "amr_hip/ndtree.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>

namespace
{

int failures = 0;

void expect_near(double a, double b, double tol, char const* what)
{
    const bool ok = std::abs(a - b) <= tol;
    std::printf("  %s: actual %.6f expected %.6f tol %.6f %s\n", what, a, b, tol, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

template <typename Tree>
std::array<double, 2> argmax_position(Tree const& tree)
{
    const auto            u    = tree.get_interior();
    double                best = -1.0;
    std::array<double, 2> pos{};
    for (std::size_t p = 0; p < tree.size(); ++p)
    {
        const auto k = tree.get_node_index_at(p);
        for (std::size_t c = 0; c < Tree::s_interior; ++c)
        {
            const double v = u[p * Tree::s_interior + c];
            if (v > best)
            {
                best = v;
                pos  = tree.cell_center(k, c);
            }
        }
    }
    return pos;
}

template <unsigned N>
void pulse_center_movement()
{
    std::printf("PulseCenterMovement<%u>\n", N);
    using tree_t = amr::hip::ndtree<2, 7, N, 2>;
    amr::hip::amr_solver<tree_t> solver(2000, 1.0, 0.4);
    const std::array<double, 2>  start = { 0.2, 0.2 };
    auto ic = [&](auto const& coords)
    {
        double r2 = 0.0;
        for (int d = 0; d < 2; ++d) r2 += std::pow(coords[d] - start[d], 2);
        return std::exp(-r2 / 0.005);
    };
    solver.initialize(ic);
    double t = 0.0, dt = 0.0, t_target = 0.5;
    while (t < t_target)
    {
        dt = solver.advance();
        if (t + dt > t_target) dt = t_target - t;
        solver.get_tree().halo_exchange_update();
        t += dt;
    }
    auto&        tree = solver.get_tree();
    const auto   pos  = argmax_position(tree);
    const auto   cs   = tree.cell_sizes(tree.get_node_index_at(0));
    const double vel[2] = { 1.0, 0.5 };
    expect_near(pos[0], start[0] + vel[0] * t_target, cs[0], "x");
    expect_near(pos[1], start[1] + vel[1] * t_target, cs[1], "y");
}

template <unsigned N>
void pulse_center_movement_amr()
{
    std::printf("PulseCenterMovementAMR<%u>\n", N);
    using tree_t        = amr::hip::ndtree<2, 7, N, 2>;
    using patch_index_t = typename tree_t::patch_index_t;
    amr::hip::amr_solver<tree_t> solver(2000, 1.0, 0.4);
    const std::array<double, 2>  start = { 0.2, 0.2 };
    auto ic = [&](auto const& coords)
    {
        double r2 = 0.0;
        for (int d = 0; d < 2; ++d) r2 += std::pow(coords[d] - start[d], 2);
        return std::exp(-r2 / 0.005);
    };
    auto criterion = [&](patch_index_t const& idx)
    {
        const auto patch   = solver.get_tree().get_patch(idx);
        double     max_val = 0;
        for (auto v : patch) max_val = std::max(max_val, v);
        if (max_val > 0.1 && idx.level() < 5) return amr::hip::refine_status_t::Refine;
        if (max_val < 0.05 && idx.level() > 1) return amr::hip::refine_status_t::Coarsen;
        return amr::hip::refine_status_t::Stable;
    };
    solver.initialize(ic);
    solver.get_tree().halo_exchange_update();
    double t = 0.0, t_target = 0.5;
    while (t < t_target)
    {
        const double dt = solver.advance();
        solver.get_tree().halo_exchange_update();
        solver.get_tree().reconstruct_tree(criterion);
        solver.get_tree().halo_exchange_update();
        t += dt;
    }
    auto&        tree   = solver.get_tree();
    const auto   pos    = argmax_position(tree);
    const auto   cs     = tree.cell_sizes(tree.get_node_index_at(0));
    const double vel[2] = { 1.0, 0.5 };
    std::printf("  leaves at end: %zu\n", tree.size());
    expect_near(pos[0], start[0] + vel[0] * t_target, cs[0], "x");
    expect_near(pos[1], start[1] + vel[1] * t_target, cs[1], "y");
}

} // namespace

int main()
{
    try
    {
        pulse_center_movement<10>();
        pulse_center_movement<50>();
        pulse_center_movement_amr<10>();
        pulse_center_movement_amr<50>();
    }
    catch (std::exception const& e)
    {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
