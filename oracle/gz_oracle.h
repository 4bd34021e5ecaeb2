/* CPU oracle for the tree-gravity path: a plain-C restatement of the reference shared-memory tree
 * (shmem/tree.h, shmem/tree.cc). TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark use it as the checker. The product library never links it.
 *
 * Pinning (each function is checked by something that does not share its failure modes):
 *   gzo_morton            against the reference's morton (oracle/_ref) and an independent
 *                         bit-by-bit interleave (system/peano.cc:317-328 has the same bit order)
 *   gzo_build             topology, orderbuf, leaf_of and every node moment compared bit for bit
 *                         with the reference build (oracle/_ref) in canonical preorder, plus the
 *                         audit invariants (tree.cc:973-1119) and root mass == sum(m)
 *   gzo_spline_force_over_r  against the reference kernel, continuity at u=1/2 and u=1, and the
 *                         Newtonian limit for r >= h
 *   gzo_open_node         against the reference open_node on random arguments
 *   gzo_accel_grouped     bit for bit against the reference accel_grouped (both regimes), and
 *                         theta -> 0 against gzo_accel_brute
 *   gzo_accel             bit for bit against the reference accel
 *   gzo_accel_grouped_tj  tidal tensor and jerk bit for bit against the reference, and against
 *                         finite differences of the direct-sum acceleration (in space and
 *                         along ballistic motion)
 *   gzo_ngb_search        lists equal to the reference's ngb_search (same order) and, as sets,
 *                         to a brute-force distance test
 *   gzo_density           bit for bit against the reference's shmem::density (h, rho, neighbour
 *                         count, iterations) in 1, 2 and 3 dimensions, with and without gas-only
 *                         neighbours and condition numbers; the kernel normalised (its integral
 *                         over the ball is 1) and dW/dh equal to finite differences of W; the
 *                         converged h meeting its own N_eff condition
 *   gzo_init_hmax, gzo_raise_hmax, gzo_ngb_search_reverse
 *                         hmax bit for bit against the reference's init_hmax and against the
 *                         maximum of the rounded-up h below each node; lists equal to the
 *                         reference's ngb_search_reverse (same order) and, as sets, to a
 *                         brute-force test of h_q <= r < h_j
 *   gzo_potential         bit for bit against the reference potential, and theta -> 0 against
 *                         the direct sum of the spline potential
 *   gzo_accel_brute       bit for bit against the reference accel_brute, Newton's third law, and
 *                         a two-body closed form
 *   gzo_audit_tree        zero violations wherever the reference's audit_tree finds none (the same
 *                         particle sets), and each invariant caught when a mistake is planted
 *                         in a copy of the tree arrays on the CPU
 *   gzo_direct_sum_f32    within float rounding of gzo_accel_brute, a two-body closed form in
 *                         float, and periodic images equal to explicitly shifted copies
 *   gzo_direct_sum_box    box = 0 bit for bit against gzo_accel_brute; a periodic box equals the
 *                         open sum over the nearest images placed explicitly
 *   gzo_build_vel, gzo_kick_nodes, gzo_refresh_walk_nodes, gzo_raise_soft
 *                         bit for bit against the reference (one thread) through vmax, vcom, the
 *                         incremental moments and the refreshed walk nodes; the refreshed COM
 *                         against the centre of mass of the drifted particles (ballistic motion
 *                         between kicks makes it exact); vmax against max |v| below each node
 *   gzo_mass_events, gzo_shift_nodes, gzo_raise_pad, gzo_raise_vmax
 *                         bit for bit against the reference's Tree::mass_event, shift_node,
 *                         raise_pad and raise_vmax called in the same order; the refreshed mass
 *                         and COM against the mass and centre of mass of the particle set with
 *                         the event applied (a particle added, removed or displaced); an event
 *                         and its negation cancelling; a leaf's pad against the float-rounded-up
 *                         running sum of its displacements and every node's pad against the
 *                         maximum over the leaves below it
 */
#ifndef GZ_ORACLE_H
#define GZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Particle SoA (tree.h:36-50). soft may be NULL, meaning softening 0. */
typedef struct {
    size_t n;
    const double *x, *y, *z, *m, *soft;
} gzo_particles;

uint64_t gzo_morton(uint32_t a, uint32_t b, uint32_t c);
double gzo_spline_force_over_r(double r, double h);
double gzo_spline_potential(double r, double h);
int gzo_open_node(double r_sq, double len, double l_open, double node_mass, double node_maxsoft,
                  double soft_target, double aold, double theta_sq, double adx, double ady,
                  double adz);

/* Root box + keys (tree.cc:355-418). keys has length n (0 for massless particles).
 * root[4] = {cx, cy, cz, side}. */
void gzo_keys(const gzo_particles* P, long long randomize_seed, uint64_t* keys, double* root);

/* Full build (tree.cc:350-526). Node ids are DFS preorder in octant order, which is the
 * reference's own numbering whenever its build runs serially. Returns an opaque handle. */
void* gzo_build(const gzo_particles* P, long long randomize_seed);
/* The same with the root box displaced by SHMEM_TREE_SHIFT's fractions (tree.cc:400-407), which
 * the reference reads from the environment. Not thread-safe (one build at a time). */
void* gzo_build_shift(const gzo_particles* P, long long randomize_seed, double fx, double fy,
                      double fz);
void gzo_tree_free(void* t);
/* out[0]=nnodes, out[1]=root, out[2]=n_live */
void gzo_tree_sizes(void* t, long long* out);
void gzo_tree_nodes(void* t, int* first, int* next, int* plo, int* phi, int* parent,
                    double* size, double* mass, double* cx, double* cy, double* cz,
                    double* soft, double* delta, double* gcx, double* gcy, double* gcz,
                    double* wn_s, float* wn_soft, float* wn_len);
void gzo_tree_particles(void* t, uint32_t* orderbuf, uint32_t* rank, int* leaf_of);

/* Build with particle velocities (tree.h:384-385, want_vcom = true): also per-node vmax and
 * centre-of-mass velocity, and zeroed incremental moments. */
void* gzo_build_vel(const gzo_particles* P, long long randomize_seed, const double* vx,
                    const double* vy, const double* vz);
/* Tree reuse (tree.cc:211-319; the kick feed of mfm.cc:3798-3811). */
void gzo_kick_nodes(void* t, const uint32_t* active, size_t na, const double* vx, const double* vy,
                    const double* vz, const double* m, double* vlx, double* vly, double* vlz,
                    double tau);
void gzo_refresh_walk_nodes(void* t, double tau);
void gzo_raise_soft(void* t, const uint32_t* idx, size_t n, const double* soft);
void gzo_tree_reuse(void* t, float* vmax, double* vcom, double* dp, double* dq, double* wcom,
                    double* wmass);
/* Events between rebuilds that are not a drift or a kick, each for the leaf that held particle
 * idx[k] at the build (particles not in the tree are skipped), applied in the order of k:
 * Tree::mass_event (tree.cc:239-251): a particle of mass dm[k] (negative removes one) at x, v,
 * appearing at tau since the build, with x unwrapped against each ancestor's current COM in a
 * periodic box of side box (box <= 0: open); needs a tree built with velocities.
 * Tree::shift_node (tree.cc:253-258): a particle of mass m[k] displaced by dx without a velocity
 * change; needs velocities at build. Tree::raise_pad (tree.cc:323-330): a non-ballistic
 * displacement of d[k] > 0. Tree::raise_vmax (tree.cc:276-285): the speed bound raised to
 * speed[k]. */
void gzo_mass_events(void* t, const uint32_t* idx, size_t n, const double* dm, const double* x,
                     const double* y, const double* z, const double* vx, const double* vy,
                     const double* vz, double tau, double box);
void gzo_shift_nodes(void* t, const uint32_t* idx, size_t n, const double* m, const double* dx,
                     const double* dy, const double* dz);
void gzo_raise_pad(void* t, const uint32_t* idx, size_t n, const double* d);
void gzo_raise_vmax(void* t, const uint32_t* idx, size_t n, const float* speed);
/* Per node: the displacement pad (float) and the mass change since the build (0 without
 * velocities at build); either may be NULL. */
void gzo_tree_pads(void* t, float* xpad, double* dmass);
/* sorted keys (n_live) */
void gzo_tree_skeys(void* t, uint64_t* skey);

/* Walks. counts (may be NULL) receives {nodes_visited, multipole, pairs} as SHMEM_COUNT_WALK
 * defines them (tree.cc:671-674). aold may be NULL. */
void gzo_accel_grouped(void* t, const gzo_particles* P, const uint32_t* targets, size_t nt,
                       double theta, double G, int batch, const double* aold, double* ax,
                       double* ay, double* az, long long* counts);
/* With the tidal tensor (tidal: 6 x nt, xx yy zz xy yz xz, no G) and the jerk (jerk: 3 x nt, G
 * included; needs vx/vy/vz and a tree built with velocities) (tree.cc:811-871). */
void gzo_accel_grouped_tj(void* t, const gzo_particles* P, const uint32_t* targets, size_t nt,
                          double theta, double G, int batch, const double* aold, const double* vx,
                          const double* vy, const double* vz, double* ax, double* ay,
                          double* az, double* tidal, double* jerk, long long* counts);
void gzo_accel(void* t, const gzo_particles* P, const uint32_t* targets, size_t nt, double theta,
               double G, const double* aold, double* ax, double* ay, double* az);
/* Neighbour search (hydro.cc:37-108): CSR of all particles within radius[q] of centre q, in the
 * reference's traversal order (= Morton order); found written while it fits in cap. */
size_t gzo_ngb_search(void* t, const gzo_particles* P, const double* cx, const double* cy,
                      const double* cz, const double* radius, size_t nq, double tau, size_t* start,
                      uint32_t* found, size_t cap);
/* init_hmax (tree.cc:332-338): node hmax from the gas kernel radii h[0..min(nh, n_gas)), kept in
 * the tree for the reverse search; copied to hmax_out (nnodes floats) when non-NULL. */
void gzo_init_hmax(void* t, const double* h, size_t nh, size_t n_gas, float* hmax_out);
/* The node hmax (nnodes floats; zeros before gzo_init_hmax). */
void gzo_tree_hmax(void* t, float* out);
/* raise_hmax (tree.cc:293-301) for the leaf of each listed particle i, to h[i] (by particle). */
void gzo_raise_hmax(void* t, const uint32_t* idx, size_t n, const double* h);
/* ngb_search_reverse (hydro.cc:110-148): CSR of the gas particles j with hq[q] <= r < h[j] for
 * each centre q, in traversal (Morton) order; found written while it fits in cap. */
size_t gzo_ngb_search_reverse(void* t, const gzo_particles* P, const double* h, size_t n_gas,
                              const double* cx, const double* cy, const double* cz,
                              const double* hq, size_t nq, double tau, size_t* start,
                              uint32_t* found, size_t cap);
/* The cubic-spline hydro kernel W(r, h) and dW/dh at fixed r (hydro.h:38-65), support h. */
double gzo_kernel_w(double r, double h, int dim);
double gzo_kernel_dwdh(double r, double h, int dim);
/* shmem::density (hydro.cc:150-327), non-periodic, tree aged tau, without the neighbour cache or
 * lazy drift: for each target (a particle index) the support radius h solving
 * N_eff(h) = ball_vol(h) sum_j W(r_j, h) = des_ngb (widened by the face-closure and, with
 * cond_prev (by particle, may be NULL), the condition-number corrections) to within ngb_tol, by
 * bracket-guarded Newton from h_start[t] (NULL: the mean-density guess over the bounding box of
 * all particles), counting gas neighbours only (type NULL: all gas). Outputs h, rho = sum m_j W,
 * the neighbours inside h and the iterations taken (100: not converged). */
void gzo_density(void* t, const gzo_particles* P, const uint8_t* type, const uint32_t* targets,
                 size_t nt, double des_ngb, double ngb_tol, const double* h_start, int n_dims,
                 const double* cond_prev, double tau, double* h_out, double* rho_out,
                 int* nngb_out, int* iters_out);
void gzo_potential(void* t, const gzo_particles* P, const uint32_t* targets, size_t nt,
                   double theta, double G, const double* aold, double* phi);
void gzo_accel_brute(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                     double* ax, double* ay, double* az);
/* The reference's device direct sum (grav_cuda.cu:65-141) in its default float arithmetic, and
 * accel_brute's FP64 form, both with an optional periodic box (box = 0: open). */
void gzo_direct_sum_f32(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                        double box, double* ax, double* ay, double* az);
void gzo_direct_sum_box(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                        double box, double* ax, double* ay, double* az);

/* A tree as plain host arrays (node ids any numbering, children threaded by first/next as in
 * tree.cc:340-348), for gzo_audit_tree. vmax and xpad may be NULL (zero). */
typedef struct {
    size_t nnodes, n_live;
    int root;
    const int *first, *next, *parent, *plo, *phi;
    const double *mass, *cx, *cy, *cz; /* walk-node mass and centre of mass at tau       */
    const double *size, *gcx, *gcy, *gcz; /* geometric side and centre                 */
    const float* wn_soft;                 /* node softening bound                        */
    const float *vmax, *xpad;
    double tau;                           /* time since the build                        */
    const uint32_t* orderbuf;
    const int* leaf_of;
} gzo_tree_view;
/* audit_tree (tree.cc:973-1119) without h, box or lazy drift: the number of violated invariants,
 * the first described in msg. */
long long gzo_audit_tree(const gzo_tree_view* T, const gzo_particles* P, const double* vx,
                         const double* vy, const double* vz, int check_com, char* msg,
                         size_t msglen);

#ifdef __cplusplus
}
#endif
#endif
