// test_slots.cpp -- the reference's fixture scenes written against the C++ slot
// API mirror (scopi_amd/host), run on the GPU; prints one JSON object per case
// with the final dynamic state, which tests/test_gpu_host_api.py compares with
// test/references/*.json.  Scene set-ups follow test/test_sphere.cpp:341-484,
// test/test_obstacles.cpp:216-254, test/test_gradient.cpp:156-213, 377-439 and
// test/test_contacts_kdtree.cpp:14-72.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>

#include <scopi/objects/types/plane.hpp>
#include <scopi/objects/types/sphere.hpp>
#include <scopi/objects/types/superellipsoid.hpp>
#include <scopi/solver.hpp>
#include <scopi/vap/vap_fpd.hpp>

using namespace scopi;

// Output files as the reference's tests ask for them (test_sphere.cpp:369-378): one at the
// start and one after the last step, {path}/{filename}_{step:04d}.json; the printed line
// names the last one.
template <class S>
std::string reference_output(S& solver, const char* path, const char* filename, std::size_t total_it) {
    auto params = solver.get_params();
    params.solver_params.output_frequency = total_it;
    params.solver_params.filename = filename;
    params.solver_params.path = path;
    char name[256];
    std::snprintf(name, sizeof(name), "%s/%s_%04zu.json", path, filename, total_it);
    return name;
}

template <std::size_t dim>
void dump(const char* name, scopi_container<dim>& p, const std::string& file) {
    std::printf("{\"case\": \"%s\", \"output\": \"%s\", \"objects\": [", name, file.c_str());
    for (std::size_t k = 0; k < p.size(); ++k) {
        std::printf("%s{\"position\": [%.17g, %.17g], \"quaternion\": [%.17g, %.17g, %.17g, %.17g], "
                    "\"velocity\": [%.17g, %.17g], \"omega\": %.17g}",
                    k ? ", " : "", p.pos()[k][0], p.pos()[k][1], p.q()[k][0], p.q()[k][1], p.q()[k][2], p.q()[k][3],
                    p.v()[k][0], p.v()[k][1], p.omega()[k]);
    }
    std::printf("]}\n");
    std::fflush(stdout);
}

template <class SolverType>
void two_spheres(bool asym) {
    constexpr std::size_t dim = 2;
    double dt = .005;
    std::size_t total_it = 1000;
    double dy = asym ? 0.05 : 0.0;
    sphere<dim> s1({{-0.2, -dy}}, 0.1);
    sphere<dim> s2({{0.2, dy}}, 0.1);
    auto p = property<dim>().mass(1.).moment_inertia(0.1);
    scopi_container<dim> particles;
    particles.push_back(s1, p.desired_velocity({{0.25, 0}}));
    particles.push_back(s2, p.desired_velocity({{-0.25, 0}}));
    SolverType solver(particles);
    const char* name = asym ? "two_spheres_asymmetrical" : "two_spheres_symmetrical";
    const auto file = reference_output(solver, "two_spheres", name, total_it);
    solver.run(dt, total_it);
    dump(name, particles, file);
}

template <class SolverType>
void critical_2d_spheres() {
    constexpr std::size_t dim = 2;
    double dt = .01;
    std::size_t total_it = 20;
    scopi_container<dim> particles;
    int n = 3;
    std::minstd_rand0 generator(123);
    std::uniform_real_distribution<double> distrib_r(0.2, 0.4);
    std::uniform_real_distribution<double> distrib_move_x(-0.1, 0.1);
    std::uniform_real_distribution<double> distrib_move_y(-0.1, 0.1);
    std::uniform_real_distribution<double> distrib_velocity(2., 5.);
    auto prop = property<dim>().mass(1.).moment_inertia(0.1);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            double r = distrib_r(generator);
            double x = (i + 0.5) + distrib_move_x(generator);
            double y = (j + 0.5) + distrib_move_y(generator);
            double velocity = distrib_velocity(generator);
            sphere<dim> s1({{x, y}}, r);
            particles.push_back(s1, prop.desired_velocity({{velocity, 0.}}));
            r = distrib_r(generator);
            x = (n + i + 0.5) + distrib_move_x(generator);
            y = (j + 0.5) + distrib_move_y(generator);
            velocity = distrib_velocity(generator);
            sphere<dim> s2({{x, y}}, r);
            particles.push_back(s2, prop.desired_velocity({{-velocity, 0.}}));
        }
    }
    SolverType solver(particles);
    const auto file = reference_output(solver, "critical_2d_spheres", "2d_case_spheres", total_it);
    solver.run(dt, total_it);
    dump("2d_case_spheres", particles, file);
}

void sphere_sphere_moving() {
    constexpr std::size_t dim = 2;
    double dt = .005;
    std::size_t total_it = 100;
    sphere<dim> obstacle({{0., 0.}}, 1.);
    sphere<dim> s({{1., std::sqrt(3.)}}, 1.);
    auto prop = property<dim>().mass(1.).moment_inertia(0.1);
    scopi_container<dim> particles;
    particles.push_back(obstacle, property<dim>().deactivate());
    particles.push_back(s, prop.force({{0., -10.}}));
    ScopiSolver<dim, NoFriction, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    const auto file = reference_output(solver, "obstacles", "obstacles_sphere_sphere_moving", total_it);
    solver.run(dt, total_it);
    dump("obstacles_sphere_sphere_moving", particles, file);
}

void three_spheres() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    sphere<dim> s1({{-1.7, 1.3}}, 0.5);
    sphere<dim> s2({{0.5, 1.7}}, 0.5);
    sphere<dim> s3({{4.5, 1.3}}, 0.5);
    particles.push_back(s1, property<dim>().mass(1).moment_inertia(1).force({{1.7, -1.3}}));
    particles.push_back(s2, property<dim>().mass(1).moment_inertia(1).force({{-0.5, -1.7}}));
    particles.push_back(s3, property<dim>().mass(1).moment_inertia(1).force({{-4.5, -1.3}}));
    double Tf = 6, dt = 0.1;
    ScopiSolver<dim, NoFriction, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.1;
    params.optim_params.dynamic_descent = true;
    std::size_t total_it = Tf / dt;
    const auto file = reference_output(solver, "three_spheres", "3spheres_nofriction", total_it);
    solver.run(dt, total_it);
    dump("3spheres_nofriction", particles, file);
}

// test/test_gradient.cpp:505-572: the three spheres under FrictionFixedPoint, mu 0.5
void three_spheres_friction_fp() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    sphere<dim> s1({{-1.7, 1.3}}, 0.5);
    sphere<dim> s2({{0.5, 1.7}}, 0.5);
    sphere<dim> s3({{4.5, 1.3}}, 0.5);
    particles.push_back(s1, property<dim>().mass(1).moment_inertia(1).force({{1.7, -1.3}}));
    particles.push_back(s2, property<dim>().mass(1).moment_inertia(1).force({{-0.5, -1.7}}));
    particles.push_back(s3, property<dim>().mass(1).moment_inertia(1).force({{-4.5, -1.3}}));
    double Tf = 7, dt = 0.1;
    ScopiSolver<dim, FrictionFixedPoint, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.1;
    params.optim_params.dynamic_descent = true;
    params.default_contact_property.mu = 0.5;
    params.default_contact_property.fixed_point_tol = 1e-6;
    params.default_contact_property.fixed_point_max_iter = 1000;
    std::size_t total_it = Tf / dt;
    const auto file = reference_output(solver, "three_spheres_fp", "3spheres_friction_fp", total_it);
    solver.run(dt, total_it);
    dump("3spheres_friction_fp", particles, file);
}

// test/test_gradient.cpp:215-282 with mu = `mu`: a sphere on a horizontal plane under gravity
// tilted by pi/4, FrictionFixedPoint on the device time loop
void sphere_plane_fixed_point(double mu) {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    const double radius = 1., g = 1., dt = 0.01, h = 2. * radius, alpha = PI / 4;
    const std::size_t total_it = 500;
    plane<dim> pl({{0., 0.}}, PI / 2);
    sphere<dim> s({{0, h}}, radius);
    particles.push_back(pl, property<dim>().deactivate());
    particles.push_back(s, property<dim>().mass(1).moment_inertia(1. * radius * radius / 2.).force(
                               {{g * std::sin(alpha), -g * std::cos(alpha)}}));
    ScopiSolver<dim, FrictionFixedPoint, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.01;
    params.optim_params.dynamic_descent = true;
    params.default_contact_property.mu = mu;
    params.default_contact_property.fixed_point_tol = 1e-6;
    params.default_contact_property.fixed_point_max_iter = 1000;
    params.solver_params.output_frequency = std::size_t(-1);
    solver.run(dt, total_it);
    char name[64];
    std::snprintf(name, sizeof(name), "sphere_plane_fp_mu%g", mu);
    std::printf("{\"case\": \"%s\", \"outer_iterations\": %lld, \"objects\": [", name,
                (long long)solver.last_info().outer_iterations);
    std::printf("{\"position\": [%.17g, %.17g], \"quaternion\": [%.17g, %.17g, %.17g, %.17g], "
                "\"velocity\": [%.17g, %.17g], \"omega\": %.17g}]}\n",
                particles.pos()[1][0], particles.pos()[1][1], particles.q()[1][0], particles.q()[1][1],
                particles.q()[1][2], particles.q()[1][3], particles.v()[1][0], particles.v()[1][1], particles.omega()[1]);
    std::fflush(stdout);
}

// The contact slot and the optim slot called one by one with FrictionFixedPoint: a unit sphere
// (m 1, J 1/2) touching a horizontal plane with velocity (1, -1); the contact's own mu (0.2) and
// fixed-point controls drive the solve.  Coulomb's answer for one step: the normal impulse
// cancels the approach, the friction impulse is mu times it (the contact point still slides),
// so u = (1 - mu, 0), omega = -mu r / J = -0.4.
void fixed_point_slot() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    particles.push_back(plane<dim>({{0., 0.}}, PI / 2), property<dim>().deactivate());
    particles.push_back(sphere<dim>({{0., 1.}}, 1.), property<dim>().mass(1).moment_inertia(0.5).velocity({{1., -1.}}));
    ContactsParams<contact_kdtree<FrictionFixedPoint>> cp;
    contact_kdtree<FrictionFixedPoint> cont(cp);
    auto contacts = cont.run(particles);
    for (auto& c : contacts) {
        c.property.mu = 0.2;
        c.property.fixed_point_tol = 1e-10;
        c.property.fixed_point_max_iter = 1000;
    }
    OptimGradient<apgd> optim;
    optim.set_timestep(0.01);
    optim.get_params().tolerance = 1e-11;
    optim.get_params().max_ite = 100000;
    optim.run(particles, contacts, 0);
    const auto& u = optim.get_uadapt();
    const auto& w = optim.get_wadapt();
    std::printf("{\"case\": \"fixed_point_slot\", \"n\": %zu, \"u\": [%.17g, %.17g], \"omega\": %.17g, "
                "\"outer_iterations\": %lld}\n",
                contacts.size(), u[0], u[1], w[2], (long long)optim.last_info().outer_iterations);
    std::fflush(stdout);
}

// test/test_gradient.cpp:441-503 / 574-640: the three spheres under Viscous (defaults) or
// ViscousFriction (mu 0.5, fixed_point_tol 1e-3, gamma_min -1.4)
template <class problem_t>
void three_spheres_viscous() {
    constexpr std::size_t dim = 2;
    constexpr bool friction = std::is_same<problem_t, ViscousFriction>::value;
    scopi_container<dim> particles;
    sphere<dim> s1({{-1.7, 1.3}}, 0.5);
    sphere<dim> s2({{0.5, 1.7}}, 0.5);
    sphere<dim> s3({{4.5, 1.3}}, 0.5);
    particles.push_back(s1, property<dim>().mass(1).moment_inertia(1).force({{1.7, -1.3}}));
    particles.push_back(s2, property<dim>().mass(1).moment_inertia(1).force({{-0.5, -1.7}}));
    particles.push_back(s3, property<dim>().mass(1).moment_inertia(1).force({{-4.5, -1.3}}));
    double Tf = 7, dt = 0.1;
    ScopiSolver<dim, problem_t, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.1;
    params.optim_params.dynamic_descent = true;
    if constexpr (friction) {
        params.default_contact_property.mu = 0.5;
        params.default_contact_property.fixed_point_tol = 1e-3;
        params.default_contact_property.fixed_point_max_iter = 1000;
        params.default_contact_property.gamma = 0;
        params.default_contact_property.gamma_min = -1.4;
        params.default_contact_property.gamma_tol = 1e-6;
    }
    std::size_t total_it = Tf / dt;
    const char* name = friction ? "3spheres_viscous_friction" : "3spheres_viscous";
    const auto file = reference_output(solver, "three_spheres_viscous", name, total_it);
    solver.run(dt, total_it);
    dump(name, particles, file);
}

// test/test_gradient.cpp:86-155 / 284-373: a sphere dropped on a plane under tilted gravity, the
// gravity reversed between two runs of the same solver (the contacts' gamma carries over)
template <class problem_t>
void sphere_plane_reversed_gravity() {
    constexpr std::size_t dim = 2;
    constexpr bool friction = std::is_same<problem_t, ViscousFriction>::value;
    scopi_container<dim> particles;
    const double radius = 1., g = 1., dt = 0.01, h = 2. * radius, alpha = PI / 4;
    plane<dim> pl({{0., 0.}}, PI / 2);
    sphere<dim> s({{0, h}}, radius);
    particles.push_back(pl, property<dim>().deactivate());
    particles.push_back(s, property<dim>().mass(1).moment_inertia(1. * radius * radius / 2.).force(
                               {{g * std::sin(alpha), -g * std::cos(alpha)}}));
    ScopiSolver<dim, problem_t, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.1;
    params.optim_params.dynamic_descent = true;
    params.solver_params.output_frequency = std::size_t(-1);
    if constexpr (friction) {
        params.default_contact_property.mu = 0.5;
        params.default_contact_property.fixed_point_tol = 1e-3;
        params.default_contact_property.fixed_point_max_iter = 1000;
        params.default_contact_property.gamma = 0;
        params.default_contact_property.gamma_min = -1.4;
        params.default_contact_property.gamma_tol = 1e-6;
    }
    solver.run(dt, 500);
    particles.f()[1] = {{-g * std::sin(alpha), g * std::cos(alpha)}};
    solver.run(dt, friction ? 100 : 500);
    std::printf("{\"case\": \"%s\", \"objects\": [", friction ? "reversed_viscous_friction" : "reversed_viscous");
    std::printf("{\"position\": [%.17g, %.17g], \"quaternion\": [%.17g, %.17g, %.17g, %.17g], "
                "\"velocity\": [%.17g, %.17g], \"omega\": %.17g}]}\n",
                particles.pos()[1][0], particles.pos()[1][1], particles.q()[1][0], particles.q()[1][1],
                particles.q()[1][2], particles.q()[1][3], particles.v()[1][0], particles.v()[1][1], particles.omega()[1]);
    std::fflush(stdout);
}

// The contact slot, the optim slot and update_contact_properties called one by one with Viscous:
// a unit sphere touching a plane, pulled away at speed 1 (dt 0.1), its contact at gamma = -2 holds it
// (u = 0) and gamma rises by the impulse to -1.
void viscous_slot() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    particles.push_back(plane<dim>({{0., 0.}}, PI / 2), property<dim>().deactivate());
    particles.push_back(sphere<dim>({{0., 1.}}, 1.), property<dim>().mass(1).moment_inertia(0.5).velocity({{0., 1.}}));
    ContactsParams<contact_kdtree<Viscous>> cp;
    contact_kdtree<Viscous> cont(cp);
    auto contacts = cont.run(particles);
    for (auto& c : contacts) c.property.gamma = -2.;
    OptimGradient<apgd> optim;
    optim.set_timestep(0.1);
    optim.get_params().tolerance = 1e-12;
    optim.get_params().max_ite = 100000;
    optim.extra_steps_before_solve(contacts);
    while (optim.should_solve()) {
        optim.run(particles, contacts, 0);
        optim.extra_steps_after_solve(contacts, particles);
    }
    optim.update_contact_properties(contacts);
    const auto& u = optim.get_uadapt();
    std::printf("{\"case\": \"viscous_slot\", \"n\": %zu, \"u\": [%.17g, %.17g], \"gamma\": %.17g, \"rows\": %zu}\n",
                contacts.size(), u[0], u[1], contacts[0].property.gamma, optim.lagrange_multiplier().size());
    std::fflush(stdout);
}

void sphere_plane_friction() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    plane<dim> pl({{0., 0.}}, PI / 2 - PI / 6);
    double dd = 1, rr = 2;
    sphere<dim> s({{0, (rr + dd) / std::cos(PI / 6)}}, rr);
    particles.push_back(pl, property<dim>().deactivate());
    particles.push_back(s, property<dim>().mass(1).moment_inertia(0.5 * rr * rr).force({{0, -1}}));
    double Tf = 20, dt = 0.1;
    ScopiSolver<dim, Friction, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-7;
    params.optim_params.max_ite = 10000;
    params.optim_params.alpha = 0.1;
    params.optim_params.dynamic_descent = true;
    params.default_contact_property.mu = 0.1;
    std::size_t total_it = Tf / dt;
    const auto file = reference_output(solver, "friction", "sphere_plane_friction_mu01", total_it);
    solver.run(dt, total_it);
    dump("sphere_plane_friction_mu01", particles, file);
}

// A periodic box through ScopiSolver(box, particles): two spheres moving towards each other
// across the periodic x bound of [0, 1] x [0, 1] (BoxDomain::with_periodicity(0)).
void periodic_pair() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    sphere<dim> s1({{0.012, 0.5}}, 0.01);
    sphere<dim> s2({{0.985, 0.503}}, 0.012);
    particles.push_back(s1, property<dim>().mass(1.).moment_inertia(0.1).desired_velocity({{-0.5, 0.}}));
    particles.push_back(s2, property<dim>().mass(1.).moment_inertia(0.1).desired_velocity({{0.5, 0.}}));
    BoxDomain<dim> box({{0., 0.}}, {{1., 1.}});
    box.with_periodicity(0);
    ScopiSolver<dim, NoFriction, OptimGradient<apgd>, contact_kdtree, vap_fixed> solver(box, particles);
    auto params = solver.get_params();
    params.contact_method_params.dmax = 0.02;
    params.contact_method_params.kd_tree_radius = 0.1 * 0.1;
    const auto file = reference_output(solver, "periodic", "periodic_pair", 40);
    solver.run(0.005, 40);
    dump("periodic_pair", particles, file);
}

// test_contacts_kdtree.cpp:14-72 through the contact slot alone, then the optim slot alone.
void slots_one_by_one() {
    constexpr std::size_t dim = 2;
    scopi_container<dim> particles;
    sphere<dim> s1({{0., 0.}}, 0.1);
    sphere<dim> s2({{1., 1.}}, 0.2);
    sphere<dim> s3({{5., 10.}}, 0.1);
    particles.push_back(s1, property<dim>().velocity({{1.0, 1.0}}).moment_inertia(0.1));
    particles.push_back(s2, property<dim>().moment_inertia(0.1));
    particles.push_back(s3, property<dim>().moment_inertia(0.1));
    ContactsParams<contact_kdtree<NoFriction>> params;
    params.kd_tree_radius = 100.;
    contact_kdtree<NoFriction> cont(params);
    auto contacts = cont.run(particles);
    std::printf("{\"case\": \"contacts_kdtree\", \"n\": %zu, \"i\": %zu, \"j\": %zu, \"dij\": %.17g, "
                "\"nij\": [%.17g, %.17g], \"pi\": [%.17g, %.17g], \"pj\": [%.17g, %.17g]}\n",
                contacts.size(), contacts[0].i, contacts[0].j, contacts[0].dij, contacts[0].nij[0], contacts[0].nij[1],
                contacts[0].pi[0], contacts[0].pi[1], contacts[0].pj[0], contacts[0].pj[1]);
    OptimGradient<apgd> optim;
    optim.set_timestep(1.0);
    optim.get_params().tolerance = 1e-10;
    optim.run(particles, contacts, 0);
    const auto& u = optim.get_uadapt();
    std::printf("{\"case\": \"optim_slot\", \"lambda\": %.17g, \"u0\": [%.17g, %.17g], \"u1\": [%.17g, %.17g]}\n",
                optim.lagrange_multiplier()[0], u[0], u[1], u[3], u[4]);
    std::fflush(stdout);
}

// test_closest_points.cpp:1632-1700 (superellipsoid_superellipsoid_3d and its 30-degree variant)
// through the contact slot.
void superellipsoids_3d() {
    constexpr std::size_t dim = 3;
    const double PI = 3.14159265358979323846, c30 = std::sqrt(3.) / 2., s30 = 0.5, dist = 0.4;
    for (int rot = 0; rot < 2; ++rot) {
        scopi_container<dim> particles;
        auto p = property<dim>().mass(1.).moment_inertia({{0.1, 0.1, 0.1}});
        if (rot == 0) {
            particles.push_back(superellipsoid<dim>({{-0.2, 0.0, 0.0}}, {quaternion(0.)}, {{0.1, 0.2, 0.3}}, {{1, 1}}), p);
            particles.push_back(superellipsoid<dim>({{0.2, 0.0, 0.0}}, {quaternion(0.)}, {{0.1, 0.2, 0.3}}, {{1, 1}}), p);
        } else {
            particles.push_back(superellipsoid<dim>({{-dist * c30, dist * s30, 0.}}, {quaternion(-PI / 6.)},
                                                    {{0.1, 0.4, 0.5}}, {{1, 1}}), p);
            particles.push_back(superellipsoid<dim>({{dist * c30, -dist * s30, 0.}}, {quaternion(PI - PI / 6.)},
                                                    {{0.2, 0.3, 0.6}}, {{1, 1}}), p);
        }
        ContactsParams<contact_kdtree<NoFriction>> params;
        params.dmax = 1.0;
        contact_kdtree<NoFriction> cont(params);
        auto contacts = cont.run(particles);
        std::printf("{\"case\": \"superellipsoids_3d_%d\", \"n\": %zu, \"dij\": %.17g, \"nij\": [%.17g, %.17g, %.17g], "
                    "\"pi\": [%.17g, %.17g, %.17g], \"pj\": [%.17g, %.17g, %.17g]}\n",
                    rot, contacts.size(), contacts[0].dij, contacts[0].nij[0], contacts[0].nij[1], contacts[0].nij[2],
                    contacts[0].pi[0], contacts[0].pi[1], contacts[0].pi[2], contacts[0].pj[0], contacts[0].pj[1],
                    contacts[0].pj[2]);
    }
    std::fflush(stdout);
}

// test_closest_points.cpp:1502-1562 (superellipsoid_superellipsoid_2d and its 30-degree variant)
// through the contact slot, 2D.
void superellipsoids_2d() {
    constexpr std::size_t dim = 2;
    const double PI = 3.14159265358979323846, c30 = std::sqrt(3.) / 2., s30 = 0.5, dist = 0.4;
    for (int rot = 0; rot < 2; ++rot) {
        scopi_container<dim> particles;
        auto p = property<dim>().mass(1.).moment_inertia(0.1);
        if (rot == 0) {
            particles.push_back(superellipsoid<dim>({{-0.2, 0.0}}, {quaternion(0.)}, {{0.1, 0.2}}, {{1}}), p);
            particles.push_back(superellipsoid<dim>({{0.2, 0.0}}, {quaternion(0.)}, {{0.1, 0.3}}, {{1}}), p);
        } else {
            particles.push_back(
                superellipsoid<dim>({{-dist * c30, dist * s30}}, {quaternion(-PI / 6.)}, {{0.1, 0.4}}, {{1}}), p);
            particles.push_back(
                superellipsoid<dim>({{dist * c30, -dist * s30}}, {quaternion(PI - PI / 6.)}, {{0.2, 0.3}}, {{1}}), p);
        }
        ContactsParams<contact_kdtree<NoFriction>> params;
        params.dmax = 1.0;
        contact_kdtree<NoFriction> cont(params);
        auto contacts = cont.run(particles);
        std::printf("{\"case\": \"superellipsoids_2d_%d\", \"n\": %zu, \"dij\": %.17g, \"nij\": [%.17g, %.17g, 0], "
                    "\"pi\": [%.17g, %.17g, 0], \"pj\": [%.17g, %.17g, 0]}\n",
                    rot, contacts.size(), contacts[0].dij, contacts[0].nij[0], contacts[0].nij[1], contacts[0].pi[0],
                    contacts[0].pi[1], contacts[0].pj[0], contacts[0].pj[1]);
    }
    std::fflush(stdout);
}

// Object states after a run: position, velocity (and the 3D rotation velocity) of every object.
template <std::size_t dim>
void dump_state(const char* name, scopi_container<dim>& p, std::size_t steps) {
    std::printf("{\"case\": \"%s\", \"steps\": %zu, \"pos\": [", name, steps);
    for (std::size_t k = 0; k < p.size(); ++k)
        for (std::size_t d = 0; d < dim; ++d) std::printf("%s%.17g", (k || d) ? ", " : "", p.pos()[k][d]);
    std::printf("], \"v\": [");
    for (std::size_t k = 0; k < p.size(); ++k)
        for (std::size_t d = 0; d < dim; ++d) std::printf("%s%.17g", (k || d) ? ", " : "", p.v()[k][d]);
    std::printf("]}\n");
    std::fflush(stdout);
}

// demos/box_spheres_3d.cpp:10-112 written against the slot mirror, run for total_it steps (the demo
// runs 1000): the demo's scene, draws, solver slots and parameters; its xtensor vectors are the
// mirror's fixed arrays.  The solver writes its output files every step, as the demo's does.
void box_spheres_3d_demo(std::size_t n, std::size_t total_it) {
    constexpr std::size_t dim = 3;
    const double PI = 3.14159265358979323846;
    double width_box = 10.;
    double g = 1.;
    double r0 = width_box / n / 2.;
    double dt = 0.2 * 0.9 * r0 / (2. * g);
    double rho = 0.2 / (dt * dt);
    scopi_container<dim> particles;
    auto prop = property<dim>().force({{0., -g, 0.}});
    const fixed<3> axes_y{{0., 1., 0.}};
    const fixed<3> axes_z{{0., 0., 1.}};
    plane<dim> p_left({{0., 0., 0.}}, {quaternion(0., axes_z)});
    particles.push_back(p_left, property<dim>().deactivate());
    plane<dim> p_right({{width_box + 2 * r0, 0., 0.}}, {quaternion(0., axes_z)});
    particles.push_back(p_right, property<dim>().deactivate());
    plane<dim> p_horizontal({{0., 0., 0.}}, {quaternion(PI / 2., axes_z)});
    particles.push_back(p_horizontal, property<dim>().deactivate());
    plane<dim> p_front({{0., 0., 0.}}, {quaternion(PI / 2., axes_y)});
    particles.push_back(p_front, property<dim>().deactivate());
    plane<dim> p_back({{0., 0., width_box + 2 * r0}}, {quaternion(PI / 2., axes_y)});
    particles.push_back(p_back, property<dim>().deactivate());
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib_r(0.8 * r0, 0.9 * r0);
    std::uniform_real_distribution<double> distrib_m(1., 2.);
    std::uniform_real_distribution<double> distrib_mouve_center(-0.05, 0.05);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t k = 0; k < n; ++k) {
                double r = distrib_r(generator);
                double m = distrib_m(generator);
                double dx = distrib_mouve_center(generator);
                double dy = distrib_mouve_center(generator);
                double dz = distrib_mouve_center(generator);
                sphere<dim> s({{r0 + 0.1 + i * 2. * r0 + dx, 1.1 * r0 + 2.1 * r0 * j + dy, r0 + 0.1 + k * 2. * r0 + dz}},
                              r);
                particles.push_back(s, prop.mass(m).moment_inertia({{m * r * r / 2., m * r * r / 2., m * r * r / 2.}}));
            }
    dump_state("box_spheres_3d_initial", particles, 0);
    ScopiSolver<dim, NoFriction, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.optim_params.tolerance = 1e-3;
    params.optim_params.alpha = rho;
    params.contact_method_params.dmax = 0.9 * r0;
    params.contact_method_params.kd_tree_radius = params.contact_method_params.dmax + 2. * 0.9 * r0;
    solver.run(dt, total_it);
    dump_state("box_spheres_3d", particles, total_it);
}

// demos/segment.cpp:8-60 written against the slot mirror (default slots: NoFriction, APGD,
// kd-tree, vap_fixed), run for total_it steps (the demo runs 1000): two segment obstacles built from
// their end points, n_parts spheres from std::default_random_engine, dmax 2 dt, kd_tree_radius
// 4 max_radius.  The demo's command-line options keep their defaults.
void segment_demo(std::size_t n_parts, std::size_t total_it) {
    constexpr std::size_t dim = 2;
    double dt = .001;
    const double max_radius = 0.06;
    scopi_container<dim> particles;
    ScopiSolver<dim> solver(particles);
    segment<dim> seg1(type::position_t<dim>{0., 1.}, type::position_t<dim>{0.4, 1.});
    segment<dim> seg2(type::position_t<dim>{0.6, 1.}, type::position_t<dim>{1., 1.});
    particles.push_back(seg1, property<dim>().deactivate());
    particles.push_back(seg2, property<dim>().deactivate());
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib_x(0.1, 0.9);
    std::uniform_real_distribution<double> distrib_y(0, 0.7);
    std::uniform_real_distribution<double> distrib_r(0.01, max_radius);
    for (std::size_t i = 0; i < n_parts; ++i) {
        auto x = distrib_x(generator);
        auto y = distrib_y(generator);
        auto radius = distrib_r(generator);
        auto prop = property<dim>().desired_velocity({{0.5 - x, 2 - y}}).mass(1.).moment_inertia(0.1);
        particles.push_back(sphere<dim>({{x, y}}, radius), prop);
    }
    dump_state("segment_initial", particles, 0);
    auto params = solver.get_params();
    params.contact_method_params.dmax = 2 * dt;
    params.contact_method_params.kd_tree_radius = 4 * max_radius;
    solver.run(dt, total_it);
    dump_state("segment", particles, total_it);
}

// demos/sphere_plane.cpp:8-50 written against the slot mirror (2D): a sphere on a 30-degree incline,
// FrictionFixedPoint mu 0.1, vap_fpd, dt 0.05, total_it steps (the demo runs 200).
void sphere_plane_demo(std::size_t total_it) {
    constexpr std::size_t dim = 2;
    double radius = 1.;
    double g = 1.;
    double mass = 1.;
    double h = 2. * radius;
    auto prop = property<dim>().mass(mass).moment_inertia(mass * radius * radius / 2.);
    double dt = 0.05;
    double alpha = PI / 6.;
    scopi_container<dim> particles;
    plane<dim> p({{0., 0.}}, PI / 2. - alpha);
    sphere<dim> s({{h * std::sin(alpha), h * std::cos(alpha)}}, radius);
    particles.push_back(p, property<dim>().deactivate());
    particles.push_back(s, prop.force({{0., -g}}));
    dump_state("sphere_plane_initial", particles, 0);
    ScopiSolver<dim, FrictionFixedPoint, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.default_contact_property.mu = 0.1;
    solver.run(dt, total_it);
    dump_state("sphere_plane", particles, total_it);
}

// demos/two_ellipsoids.cpp:7-50 written against the slot mirror (2D superellipsoids, default slots),
// dt 0.005, total_it steps (the demo runs 1000).
void two_ellipsoids_demo(std::size_t total_it) {
    constexpr std::size_t dim = 2;
    double dt = .005;
    scopi_container<dim> particles;
    superellipsoid<dim> s1({{-0.2, 0.}}, {quaternion(PI / 4)}, {{.1, .05}}, 1);
    superellipsoid<dim> s2({{0.2, 0.}}, {quaternion(-PI / 4)}, {{.1, .05}}, 1);
    auto prop1 = property<dim>().desired_velocity({{0.25, 0}}).mass(1.).moment_inertia(0.1);
    auto prop2 = property<dim>().desired_velocity({{-0.25, 0}}).mass(1.).moment_inertia(0.1);
    particles.push_back(s1, prop1);
    particles.push_back(s2, prop2);
    dump_state("two_ellipsoids_initial", particles, 0);
    ScopiSolver<dim> solver(particles);
    solver.run(dt, total_it);
    dump_state("two_ellipsoids", particles, total_it);
}

// demos/pile_of_sand_spheres.cpp:9-80 written against the slot mirror (3D, FrictionFixedPoint mu 0.1,
// vap_fpd): n = 3, masses from std::default_random_engine, dt 0.01, total_it steps (the demo runs 1000).
void pile_of_sand_demo(std::size_t total_it) {
    constexpr std::size_t dim = 3;
    double dt = 0.01;
    double width_box = 10.;
    std::size_t n = 3;
    double g = 1.;
    scopi_container<dim> particles;
    auto prop = property<dim>().force({{0., -g, 0.}});
    plane<dim> p_horizontal({{0., 0., 0.}}, PI / 2.);
    particles.push_back(p_horizontal, property<dim>().deactivate());
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib_m(1., 2.);
    double r = width_box / 2. / (n + 1);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 1; j < n; ++j)
            for (std::size_t k = 0; k < n; ++k) {
                double m = distrib_m(generator);
                sphere<dim> s({{i * 2. * r, r + j * 2. * r, k * 2. * r}}, r);
                particles.push_back(s, prop.mass(m).moment_inertia({{m * r * r / 2., m * r * r / 2., m * r * r / 2.}}));
            }
    double dec_x = 0.5 * r;
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t k = 0; k < n; ++k) {
            double m = distrib_m(generator);
            sphere<dim> s({{i * 2. * r + dec_x, r, k * 2. * r}}, r);
            particles.push_back(s, prop.mass(m).moment_inertia({{m * r * r / 2., m * r * r / 2., m * r * r / 2.}}));
        }
    dump_state("pile_of_sand_initial", particles, 0);
    ScopiSolver<dim, FrictionFixedPoint, OptimGradient<apgd>, contact_kdtree, vap_fpd> solver(particles);
    auto params = solver.get_params();
    params.default_contact_property.mu = 0.1;
    solver.run(dt, total_it);
    dump_state("pile_of_sand", particles, total_it);
}

int main(int argc, char** argv) {
    std::string only = argc > 1 ? argv[1] : "";
    using kd_fixed = ScopiSolver<2, NoFriction, OptimGradient<apgd>, contact_kdtree, vap_fixed>;
    using bf_fixed = ScopiSolver<2, NoFriction, OptimGradient<apgd>, contact_brute_force, vap_fixed>;
    try {
        if (only.empty() || only == "two_spheres") {
            two_spheres<kd_fixed>(true);
            two_spheres<bf_fixed>(false);
        }
        if (only.empty() || only == "critical") critical_2d_spheres<kd_fixed>();
        if (only.empty() || only == "moving") sphere_sphere_moving();
        if (only.empty() || only == "three") three_spheres();
        if (only.empty() || only == "friction") sphere_plane_friction();
        if (only.empty() || only == "fixed_point") {
            three_spheres_friction_fp();
            sphere_plane_fixed_point(0.5);
            sphere_plane_fixed_point(0.1);
            fixed_point_slot();
        }
        if (only.empty() || only == "viscous") {
            three_spheres_viscous<Viscous>();
            three_spheres_viscous<ViscousFriction>();
            sphere_plane_reversed_gravity<Viscous>();
            sphere_plane_reversed_gravity<ViscousFriction>();
            viscous_slot();
        }
        if (only.empty() || only == "slots") slots_one_by_one();
        if (only.empty() || only == "superellipsoids") {
            superellipsoids_3d();
            superellipsoids_2d();
        }
        if (only.empty() || only == "periodic") periodic_pair();
        // the demos run only when asked for, with their step counts:
        // demos <box steps> <segment steps> <steps of sphere_plane, two_ellipsoids and pile_of_sand>
        if (only == "demos") {
            const std::size_t box_steps = argc > 2 ? std::stoul(argv[2]) : 30;
            const std::size_t seg_steps = argc > 3 ? std::stoul(argv[3]) : 200;
            box_spheres_3d_demo(10, box_steps);
            segment_demo(50, seg_steps);
            const std::size_t more = argc > 4 ? std::stoul(argv[4]) : 200;
            sphere_plane_demo(more);
            two_ellipsoids_demo(more);
            pile_of_sand_demo(more);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_slots: %s\n", e.what());
        return 1;
    }
    return 0;
}
