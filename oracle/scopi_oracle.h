/*
 * scopi_oracle.h -- C interface of the FP64 CPU restatement of SCoPI's
 * per-time-step contact pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (scopi_amd/, libscopi_hip.so) never links or calls it.
 *
 * Array conventions (shared with include/scopi_hip.h):
 *   - every per-particle vector is stored with 3 slots (x, y, z) even in 2D;
 *   - 2D rotation velocity / moment of inertia live in slot 2 (the z axis),
 *     which is the slot the reference's 6-dof u-vector uses in 2D
 *     (include/scopi/solvers/minimization_problem.hpp:15-29);
 *   - quaternions are scalar-first (w, x, y, z) (include/scopi/quaternion.hpp:86-124);
 *   - size[3]: sphere radius in size[0]; segment length in size[0];
 *     superellipsoid semi-axes in size[0..2]; planes ignore it.
 */
#ifndef SCOPI_ORACLE_H
#define SCOPI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_SPHERE = 0, ORACLE_PLANE = 1, ORACLE_SEGMENT = 2, ORACLE_SUPERELLIPSOID = 3 };
enum { ORACLE_NOFRICTION = 0, ORACLE_FRICTION = 1, ORACLE_FRICTION_FIXED_POINT = 2, ORACLE_VISCOUS = 3,
       ORACLE_VISCOUS_FRICTION = 4 };
enum { ORACLE_CONTACT_KDTREE = 0, ORACLE_CONTACT_BRUTE_FORCE = 1 };
enum { ORACLE_VAP_FIXED = 0, ORACLE_VAP_FPD = 1, ORACLE_VAP_PROJECTION = 2 };
enum { ORACLE_METHOD_APGD = 0, ORACLE_METHOD_PGD = 1 };

typedef struct oracle_params {
    int32_t problem;         /* ORACLE_NOFRICTION / ORACLE_FRICTION            */
    int32_t contact_method;  /* kd-tree predicate set / brute force            */
    int32_t vap;             /* vap_fixed / vap_fpd                            */
    int32_t method;          /* apgd / pgd                                     */
    double dmax;             /* ContactsParams::dmax (default 2)               */
    double kd_tree_radius;   /* squared search radius (default 17)            */
    double alpha;            /* apgd_params::alpha (default 0.05)              */
    int64_t max_ite;         /* apgd_params::max_ite (default 10000)           */
    double tolerance;        /* apgd_params::tolerance (default 1e-7)          */
    int32_t dynamic_descent; /* apgd_params::dynamic_descent (default true)    */
    int32_t warm_start;      /* not in the reference (default 0): NoFriction and */
                             /* Friction solves start from the last solve's      */
                             /* multipliers of the contacts found again          */
    double mu;               /* contact_property<Friction>::mu (default 0.1)   */
    double fixed_point_tol;  /* contact_property<FixedPoint> (default 1e-6)    */
    double fixed_point_max_iter; /* (default 1000)                            */
    double gamma;            /* contact_property<Viscous>::gamma (default 0)   */
    double gamma_min;        /* (default -2)                                   */
    double gamma_tol;        /* (default 1e-6)                                 */
    double objective_perturbation; /* test hook, default 0: the line search     */
                             /* compares f(lambda+) (1 + this): a rounding-level */
                             /* change that shows how far a result depends on    */
                             /* the rounding of the acceptance test              */
} oracle_params;

typedef struct oracle_stats {
    int64_t n_contacts;   /* contacts of the last step                        */
    int64_t iterations;   /* APGD outer iterations of the last solve          */
    int64_t backtracks;   /* line-search rejections of the last solve         */
    int64_t q_applies;    /* Q-matrix applications of the last solve          */
    int64_t converged;    /* 1 if the last solve met the tolerance            */
    int64_t total_iterations; /* summed over every solve since the scene was set */
    int64_t total_q_applies;
    int64_t fixed_point_iterations; /* FrictionFixedPoint / ViscousFriction: solves of the last step */
    double t_contacts, t_vap, t_solve, t_move; /* seconds, summed over steps  */
} oracle_stats;

void* oracle_scene_new(void);
void oracle_scene_free(void* scene);
/* A copy of a scene with its contact history (the solver's m_old_contacts) and parameters. */
void* oracle_scene_clone(const void* scene);

int oracle_default_params(oracle_params* p);
int oracle_set_params(void* scene, const oracle_params* p);

/* Replace the whole particle set (obstacles first: indices [0, n_inactive)).
 * squareness: n x 2 (e, n) of superellipsoids, may be NULL (then 1, 1). */
int oracle_scene_set(void* scene, int32_t dim, int64_t n, int64_t n_inactive, const int32_t* shape,
                     const double* pos, const double* q, const double* size, const double* squareness, const double* v,
                     const double* vd, const double* omega, const double* domega, const double* f,
                     const double* m, const double* J);
/* BoxDomain of the scene (dim entries each): lower / upper corner, periodic axes. */
int oracle_set_box(void* scene, const double* lo, const double* hi, const int32_t* periodic);
/* vap_projection::set_u_w: u, w are Na x 3. */
int oracle_set_projection_velocity(void* scene, const double* u, const double* w);
int oracle_get_state(void* scene, double* pos, double* q, double* v, double* omega);

/* ScopiSolver::run(dt, total_it, initial_iter) without output files. */
int oracle_run(void* scene, double dt, int64_t total_it, int64_t initial_iter);

/* The contact slot alone, on the current state; returns nc or <0. */
int64_t oracle_compute_contacts(void* scene);
/* Copy the current contact list (last step's, or the last compute). */
int oracle_get_contacts(void* scene, int64_t* ij, double* dij, double* nij, double* pi, double* pj);
int64_t oracle_n_contacts(void* scene);

/* Matrix-free operators on the current contact list. */
int oracle_amat(void* scene, const double* u, double* out);  /* 6*Na -> 3*nc */
int oracle_atmat(void* scene, const double* f, double* out); /* 3*nc -> 6*Na */

/* The optim slot alone: solve on the current contacts with the current
 * velocities taken as a-priori velocities.  lambda (size nc or 3nc), u (Na*3),
 * w (Na*3) may be NULL. */
int oracle_solve(void* scene, double dt, double* lambda, double* u, double* w);
/* The gradient Q lambda + c (+ S) of the optim slot's problem on the current contacts and a-priori
 * velocities (NoFriction and Friction), at the given multipliers; both of oracle_lambda_size. */
int oracle_gradient(void* scene, double dt, const double* lambda, double* out);
/* Replace the contact list (nc records: ij as 2 per contact, dij, nij, pi, pj as 3 per contact);
 * contact properties from the parameters. */
int oracle_set_contacts(void* scene, int64_t nc, const int64_t* ij, const double* dij, const double* nij,
                        const double* pi, const double* pj);
/* OptimGradient::update_contact_properties after oracle_solve (Viscous problems: the contacts'
 * gamma from the last solve's multipliers). */
int oracle_update_contact_properties(void* scene, double dt);
/* Replace the external forces (n x 3); the contacts of the last step and their properties stay,
 * as they do when a reference test edits particles.f() between two runs. */
int oracle_set_forces(void* scene, const double* f);
/* Size of the last solve's multiplier vector (nc, 3nc, or the Viscous layouts' variable size). */
int64_t oracle_lambda_size(void* scene);
/* The current contacts' gamma (contact_property<Viscous>::gamma): get, or set before a solve. */
int oracle_get_contact_gamma(void* scene, double* gamma);
int oracle_set_contact_gamma(void* scene, const double* gamma);
/* The vap slot alone. */
int oracle_set_a_priori_velocity(void* scene, double dt);

int oracle_get_stats(void* scene, oracle_stats* st);

/* closest_points(object_i, object_j) for one pair, no dmax filter.
 * sq_i / sq_j: superellipsoid squareness (e, n), may be NULL for other shapes. */
int oracle_closest_points(int32_t dim, int32_t shape_i, const double* pos_i, const double* q_i,
                          const double* size_i, const double* sq_i, int32_t shape_j, const double* pos_j,
                          const double* q_j, const double* size_j, const double* sq_j, double* dij, double* nij,
                          double* pi, double* pj);

/* closest_points(superellipsoid<3>, superellipsoid<3>) with the solver's
 * diagnostics: u0 = grid start (a1, b1, a2, b2), u = end point, info / nfev of hybrd1
 * (nfev counts n per Jacobian, as minpack.hpp:497-499 does) and the Jacobian count. */
int oracle_se_closest_points(const double* pos_i, const double* q_i, const double* size_i, const double* sq_i,
                             const double* pos_j, const double* q_j, const double* size_j, const double* sq_j,
                             double* dij, double* nij, double* pi, double* pj, double* u0, double* u, int32_t* info,
                             int32_t* nfev, int32_t* njev);
/* closest_points of a sphere or a plane against a superellipsoid (3D), in the dispatcher's
 * order (shape tags as above), with newton_method's start u0, end point u and iteration
 * count (-1: not converged). */
int oracle_newton_closest_points(int32_t shape_i, const double* pos_i, const double* q_i, const double* size_i,
                                 const double* sq_i, int32_t shape_j, const double* pos_j, const double* q_j,
                                 const double* size_j, const double* sq_j, double* dij, double* nij, double* pi,
                                 double* pj, double* u0, double* u, int32_t* iterations);
/* Residual and Jacobian of the Newton pairs: kind 0 = (superellipsoid i, sphere j) of
 * closest_points(sphere, superellipsoid), kind 1 = (superellipsoid i, plane j). */
int oracle_newton_residual(int32_t kind, const double* pos_i, const double* q_i, const double* size_i,
                           const double* sq_i, const double* pos_j, const double* q_j, const double* size_j,
                           const double* u, double* F, double* J);
/* closest_points(superellipsoid<2>, superellipsoid<2>) with the root u (b1, b2) and hybrd1's
 * info; 2D arrays (pos, size: 2 values; q: 4). */
int oracle_se2_closest_points(const double* pos_i, const double* q_i, const double* size_i, double e_i,
                              const double* pos_j, const double* q_j, const double* size_j, double e_j, double* dij,
                              double* nij, double* pi, double* pj, double* u, int32_t* info);
/* Residual F(u) (2) and Jacobian (2x2 row-major, may be NULL) of the SE2-SE2 system. */
int oracle_se2_residual(const double* pos_i, const double* q_i, const double* size_i, double e_i,
                        const double* pos_j, const double* q_j, const double* size_j, double e_j, const double* u,
                        double* F, double* J);
/* The 4x4 linear solve of newton_method (LU with partial pivoting), A row-major. */
int oracle_lu_solve4(const double* A, const double* b, double* x);
/* Residual F(u) (4) and Jacobian (4x4 row-major, may be NULL) of the SE3-SE3 system. */
int oracle_se_residual(const double* pos_i, const double* q_i, const double* size_i, const double* sq_i,
                       const double* pos_j, const double* q_j, const double* size_j, const double* sq_j,
                       const double* u, double* F, double* J);
/* Elementary functions of the superellipsoid solve: 0 = C library (default),
 * 1 = the portable functions of se_math.hpp, which the device evaluates identically. */
int oracle_set_se_math(int32_t portable);
/* Threads of the loops the reference runs under OpenMP (contact search over the active particles,
 * vap, velocity update and move: contact_kdtree.hpp:220, vap_fpd.hpp:74, solver.hpp:446, 535); the
 * projected-gradient solve stays serial as in the reference.  Default 1.  Returns the count set. */
int oracle_set_threads(int32_t n);
/* Test-only forms of the superellipsoid solve.  MINPACK: 0 = the reference's src/minpack.cpp
 * (default); each set bit takes one ANL-80-74 form instead: 1 enorm, 2 giant, 4 dogleg's
 * scaled-gradient loop, 8 termination order and the delta start (15: ANL-80-74).  Jacobian of the
 * SE3-SE3 and SE2-SE2 systems: 0 = the reference's hand-written one (default), 1 = forward-mode
 * differentiation of F.  The counter reports dogleg scaled-gradient steps taken before hybrd's
 * first trial point, where the reference reads an uninitialised work element (0 here). */
int oracle_set_minpack_variant(int32_t mask);
int oracle_set_jacobian_ad(int32_t ad);
int64_t oracle_dogleg_fresh_wa4_count(void);
/* One dogleg call (n = 2 or 4): r packed upper triangle by rows, wa4_last the value the
 * reference's loop reads before r (hybrd's wa4[n-1]); x receives the step. */
int oracle_minpack_dogleg(int32_t n, const double* r, const double* diag, const double* qtb, double delta,
                          double wa4_last, double* x, int32_t* gradient_branch);
/* One portable elementary function: fn 0 cos, 1 sin, 2 asin, 3 log, 4 exp, 5 pow(x, y). */
int oracle_math(int32_t fn, double x, double y, double* out);
/* superellipsoid::point / normal (a, b). */
int oracle_se_point_normal(const double* pos, const double* q, const double* size, const double* sq, double a,
                           double b, double* point, double* normal);

/* quaternion(angle, axis) of src/quaternion.cpp:12-26. */
int oracle_quaternion(double angle, const double* axis, double* q);

/* Draws of the spheres of demos/segment.cpp:30-44 (std::default_random_engine). */
int oracle_gen_segment_demo(int32_t n, double max_radius, double* x, double* y, double* r);
/* demos/box_spheres_3d.cpp's draws (radius, mass, centre offsets per sphere, in the demo's loop order) */
/* n draws of uniform_real_distribution<double>(a, b) from a default std::default_random_engine */
int oracle_gen_uniform(int32_t n, double a, double b, double* out);
int oracle_gen_box_spheres_demo(int32_t n_side, double r0, double* r, double* m, double* dx, double* dy, double* dz);
/* Scene generator of test/test_sphere.cpp:423-470 (std::minstd_rand0(123)). */
int oracle_gen_2d_case_spheres(int32_t n, double* x, double* y, double* r, double* velocity);

#ifdef __cplusplus
}
#endif

#endif
