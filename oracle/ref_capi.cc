// C entry points over the REFERENCE shared-memory tree (shmem/tree.h, shmem/tree.cc), compiled
// from the reference sources where they lie (oracle/Makefile). Test infrastructure only: the
// tests and the CPU-baseline leg of the benchmark load the resulting oracle/_ref/libgzref.so to
// pin the C restatement (oracle/gz_oracle.c) and to time the reference CPU walk. Nothing in the
// product library links or loads it.
//
// Every function here is a thin adapter: it moves plain arrays into shmem::Particles and calls the
// reference function named in its comment, unchanged.

#include "tree.h"
#include "hydro.h"

#include <cstdint>
#include <cstring>
#include <vector>

using namespace shmem;

namespace {
struct RefParticles {
    Particles P;
};
struct RefTree {
    Tree T;
    BuildTimes bt;
};
}  // namespace

extern "C" {

// Particles (tree.h:36-50). soft may be NULL (empty vector => softening 0).
void* gzref_particles_new(size_t n, const double* x, const double* y, const double* z,
                          const double* m, const double* soft) {
    auto* rp = new RefParticles();
    rp->P.x.assign(x, x + n);
    rp->P.y.assign(y, y + n);
    rp->P.z.assign(z, z + n);
    rp->P.m.assign(m, m + n);
    if (soft) rp->P.soft.assign(soft, soft + n);
    return rp;
}

void gzref_particles_free(void* p) { delete static_cast<RefParticles*>(p); }

// shmem::build (tree.h:384, tree.cc:350-526). randomize_seed < 0 disables RANDOMIZE_GRAVTREE.
void* gzref_build(void* particles, long long randomize_seed) {
    auto* rp = static_cast<RefParticles*>(particles);
    auto* rt = new RefTree();
    rt->T = build(rp->P, &rt->bt, nullptr, false, randomize_seed);
    return rt;
}

void gzref_tree_free(void* t) { delete static_cast<RefTree*>(t); }

// Sizes: out[0]=nnodes, out[1]=root, out[2]=n_live (orderbuf size).
void gzref_tree_sizes(void* t, long long* out) {
    auto* rt = static_cast<RefTree*>(t);
    out[0] = (long long)rt->T.nnodes();
    out[1] = rt->T.root;
    out[2] = (long long)rt->T.orderbuf.size();
}

// BuildTimes (tree.h:371): bbox, keys, sort, recurse, links, pack, total (ms).
void gzref_tree_times(void* t, double* out7) {
    auto* rt = static_cast<RefTree*>(t);
    const BuildTimes& b = rt->bt;
    const double v[7] = {b.bbox, b.keys, b.sort, b.recurse, b.links, b.pack, b.total};
    std::memcpy(out7, v, sizeof v);
}

// Node arrays (tree.h:228-247), each of length nnodes; any pointer may be NULL.
void gzref_tree_nodes(void* t, int* first, int* next, int* plo, int* phi, int* parent,
                      double* size, double* mass, double* cx, double* cy, double* cz,
                      double* soft, double* delta, double* gcx, double* gcy, double* gcz,
                      double* wn_s, float* wn_soft, float* wn_len) {
    auto* rt = static_cast<RefTree*>(t);
    const Tree& T = rt->T;
    const size_t nn = T.nnodes();
    for (size_t i = 0; i < nn; ++i) {
        if (first) first[i] = T.first[i];
        if (next) next[i] = T.next[i];
        if (plo) plo[i] = T.plo[i];
        if (phi) phi[i] = T.phi[i];
        if (parent) parent[i] = T.parent[i];
        if (size) size[i] = T.size[i];
        if (mass) mass[i] = T.mass[i];
        if (cx) cx[i] = T.cx[i];
        if (cy) cy[i] = T.cy[i];
        if (cz) cz[i] = T.cz[i];
        if (soft) soft[i] = T.soft[i];
        if (delta) delta[i] = T.delta[i];
        if (gcx) gcx[i] = T.gcx[i];
        if (gcy) gcy[i] = T.gcy[i];
        if (gcz) gcz[i] = T.gcz[i];
        if (wn_s) wn_s[i] = T.wn[i].s;
        if (wn_soft) wn_soft[i] = T.wn[i].soft;
        if (wn_len) wn_len[i] = T.wn[i].len;
    }
}

// orderbuf (n_live), rank (n), leaf_of (n); any pointer may be NULL.
void gzref_tree_particles(void* t, uint32_t* orderbuf, uint32_t* rank, int* leaf_of) {
    auto* rt = static_cast<RefTree*>(t);
    const Tree& T = rt->T;
    if (orderbuf) std::memcpy(orderbuf, T.orderbuf.data(), T.orderbuf.size() * sizeof(uint32_t));
    if (rank) std::memcpy(rank, T.rank.data(), T.rank.size() * sizeof(uint32_t));
    if (leaf_of) std::memcpy(leaf_of, T.leaf_of.data(), T.leaf_of.size() * sizeof(int));
}

// shmem::morton (tree.h:171-173), exposed so key tests can pin the interleave itself.
uint64_t gzref_morton(uint32_t a, uint32_t b, uint32_t c) { return morton(a, b, c); }

// shmem::spline_force_over_r (tree.h:59-67) and shmem::spline_potential (tree.h:72-80).
double gzref_spline_force_over_r(double r, double h) { return spline_force_over_r(r, h); }
double gzref_spline_potential(double r, double h) { return spline_potential(r, h); }

// shmem::open_node (tree.h:109-144).
int gzref_open_node(double r_sq, double len, double l_open, double node_mass,
                    double node_maxsoft, double soft_target, double aold, double theta_sq,
                    double adx, double ady, double adz) {
    return open_node(r_sq, len, l_open, node_mass, node_maxsoft, soft_target, aold, theta_sq,
                     adx, ady, adz) ? 1 : 0;
}

// shmem::accel_grouped (tree.h:428-440, tree.cc:626-897) with tidal/lazy/jerk/sink off.
// aold may be NULL. Outputs have length nt.
void gzref_accel_grouped(void* t, void* particles, const uint32_t* targets, size_t nt,
                         double theta, double G, int batch, const double* aold,
                         double* ax, double* ay, double* az) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    std::vector<uint32_t> tg(targets, targets + nt);
    std::vector<double> vx, vy, vz;
    accel_grouped(rt->T, rp->P, tg, theta, G, batch, vx, vy, vz, nullptr, aold);
    std::memcpy(ax, vx.data(), nt * sizeof(double));
    std::memcpy(ay, vy.data(), nt * sizeof(double));
    std::memcpy(az, vz.data(), nt * sizeof(double));
}

// shmem::accel_grouped with the tidal tensor and, when vx is given, the jerk (tree.h:414-440).
// tidal: 6 x nt (xx yy zz xy yz xz, component-major); jerk: 3 x nt; either may be NULL.
void gzref_accel_grouped_tj(void* t, void* particles, const uint32_t* targets, size_t nt,
                            double theta, double G, int batch, const double* aold,
                            const double* vx, const double* vy, const double* vz, double* ax,
                            double* ay, double* az, double* tidal, double* jerk) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    std::vector<uint32_t> tg(targets, targets + nt);
    std::vector<double> ox, oy, oz;
    std::vector<SymTensor3d> tt;
    std::vector<Vec3d> jj;
    const double* vel[3] = {vx, vy, vz};
    accel_grouped(rt->T, rp->P, tg, theta, G, batch, ox, oy, oz, tidal ? &tt : nullptr, aold,
                  nullptr, jerk ? &jj : nullptr, vx ? vel : nullptr);
    std::memcpy(ax, ox.data(), nt * sizeof(double));
    std::memcpy(ay, oy.data(), nt * sizeof(double));
    std::memcpy(az, oz.data(), nt * sizeof(double));
    for (size_t i = 0; i < nt; ++i) {
        if (tidal) {
            const SymTensor3d& q = tt[i];
            tidal[i] = q[0][0]; tidal[nt + i] = q[1][1]; tidal[2 * nt + i] = q[2][2];
            tidal[3 * nt + i] = q[0][1]; tidal[4 * nt + i] = q[1][2]; tidal[5 * nt + i] = q[0][2];
        }
        if (jerk) {
            jerk[i] = jj.empty() ? 0.0 : jj[i][0];
            jerk[nt + i] = jj.empty() ? 0.0 : jj[i][1];
            jerk[2 * nt + i] = jj.empty() ? 0.0 : jj[i][2];
        }
    }
}

// shmem::accel (tree.h:392-394, tree.cc:528-582), the per-target walk.
void gzref_accel(void* t, void* particles, const uint32_t* targets, size_t nt, double theta,
                 double G, const double* aold, double* ax, double* ay, double* az) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    std::vector<uint32_t> tg(targets, targets + nt);
    std::vector<double> vx, vy, vz;
    accel(rt->T, rp->P, tg, theta, G, vx, vy, vz, aold);
    std::memcpy(ax, vx.data(), nt * sizeof(double));
    std::memcpy(ay, vy.data(), nt * sizeof(double));
    std::memcpy(az, vz.data(), nt * sizeof(double));
}

// shmem::potential (tree.h:446-447, tree.cc:899-944).
void gzref_potential(void* t, void* particles, const uint32_t* targets, size_t nt, double theta,
                     double G, const double* aold, double* phi) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    std::vector<uint32_t> tg(targets, targets + nt);
    std::vector<double> v;
    potential(rt->T, rp->P, tg, theta, G, v, aold);
    std::memcpy(phi, v.data(), nt * sizeof(double));
}

// shmem::accel_brute (tree.h:450-451, tree.cc:946-971).
void gzref_accel_brute(void* particles, const uint32_t* targets, size_t nt, double G,
                       double* ax, double* ay, double* az) {
    auto* rp = static_cast<RefParticles*>(particles);
    std::vector<uint32_t> tg(targets, targets + nt);
    std::vector<double> vx, vy, vz;
    accel_brute(rp->P, tg, G, vx, vy, vz);
    std::memcpy(ax, vx.data(), nt * sizeof(double));
    std::memcpy(ay, vy.data(), nt * sizeof(double));
    std::memcpy(az, vz.data(), nt * sizeof(double));
}

// shmem::audit_tree (tree.h:353-355, tree.cc:973-1119), gravity-only arguments.
long long gzref_audit(void* t, void* particles, char* msg, size_t msglen) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    return audit_tree(rt->T, rp->P, nullptr, 0, nullptr, nullptr, nullptr, 0.0, nullptr, msg,
                      msglen);
}

// shmem::build with velocities and want_vcom (tree.h:384-385): vmax, vcom and the incremental
// moments, as rebuild_tree does (mfm.cc:1408-1413).
void* gzref_build_vel(void* particles, long long randomize_seed, const double* vx,
                      const double* vy, const double* vz) {
    auto* rp = static_cast<RefParticles*>(particles);
    auto* rt = new RefTree();
    const double* vel[3] = {vx, vy, vz};
    rt->T = build(rp->P, &rt->bt, vel, true, randomize_seed);
    return rt;
}

// The kick feed of mfm.cc:3798-3811, run serially in the order of `active`: Tree::raise_vmax,
// then Tree::kick_node with m (v - v_last) when v_last is given, then v_last = v.
void gzref_kick_nodes(void* t, const uint32_t* active, size_t na, const double* vx,
                      const double* vy, const double* vz, const double* m, double* vlx,
                      double* vly, double* vlz, double tau) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < na; ++k) {
        const uint32_t i = active[k];
        const int leaf = T.leaf_of[i];
        if (leaf < 0) continue;
        const Vec3d v_now{vx[i], vy[i], vz[i]};
        T.raise_vmax(leaf, (float)v_now.norm());
        if (!T.dp_x.empty() && vlx)
            T.kick_node(leaf, (v_now - Vec3d{vlx[i], vly[i], vlz[i]}) * m[i], tau);
        if (vlx) { vlx[i] = vx[i]; vly[i] = vy[i]; vlz[i] = vz[i]; }
    }
}

// Tree::refresh_walk_nodes (tree.cc:260-274).
void gzref_refresh_walk_nodes(void* t, double tau) {
    static_cast<RefTree*>(t)->T.refresh_walk_nodes(tau);
}

// Tree::raise_soft (tree.cc:311-319) for each listed particle, in order.
void gzref_raise_soft(void* t, const uint32_t* idx, size_t n, const double* soft) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < n; ++k) T.raise_soft(T.leaf_of[idx[k]], soft[idx[k]]);
}

// Tree::mass_event (tree.cc:239-251) for the leaf of each listed particle, in order, with the
// event's own mass, position and velocity.
void gzref_mass_events(void* t, const uint32_t* idx, size_t n, const double* dm, const double* x,
                       const double* y, const double* z, const double* vx, const double* vy,
                       const double* vz, double tau, double box) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < n; ++k)
        T.mass_event(T.leaf_of[idx[k]], dm[k], Vec3d{x[k], y[k], z[k]},
                     Vec3d{vx[k], vy[k], vz[k]}, tau, box);
}

// Tree::shift_node (tree.cc:253-258) for the leaf of each listed particle, in order.
void gzref_shift_nodes(void* t, const uint32_t* idx, size_t n, const double* m, const double* dx,
                       const double* dy, const double* dz) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < n; ++k)
        T.shift_node(T.leaf_of[idx[k]], m[k], Vec3d{dx[k], dy[k], dz[k]});
}

// Tree::raise_pad (tree.cc:323-330) for the leaf of each listed particle, in order.
void gzref_raise_pad(void* t, const uint32_t* idx, size_t n, const double* d) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < n; ++k) T.raise_pad(T.leaf_of[idx[k]], d[k]);
}

// Tree::raise_vmax (tree.cc:276-285) for the leaf of each listed particle, in order.
void gzref_raise_vmax(void* t, const uint32_t* idx, size_t n, const float* speed) {
    Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t k = 0; k < n; ++k) {
        const int leaf = T.leaf_of[idx[k]];
        if (leaf >= 0) T.raise_vmax(leaf, speed[k]);
    }
}

// shmem::density (hydro.h:115-119, hydro.cc:150-327), non-periodic, no neighbour cache, the tree
// aged tau. type (may be NULL: all gas) is set on the particles for the call; h_start NULL means
// the reference's own default guess; cond_prev (by particle) may be NULL.
void gzref_density(void* t, void* particles, const uint8_t* type, const uint32_t* targets,
                   size_t nt, double des_ngb, double ngb_tol, const double* h_start, int n_dims,
                   const double* cond_prev, double tau, double* h, double* rho, int* nngb,
                   int* iters) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    rt->T.t_since_build = tau;
    if (type) rp->P.type.assign(type, type + rp->P.size());
    const std::vector<uint32_t> tg(targets, targets + nt);
    const std::vector<double> hs = h_start ? std::vector<double>(h_start, h_start + nt)
                                           : std::vector<double>();
    const DensityResult r = density(rt->T, rp->P, tg, des_ngb, ngb_tol, hs, 0.0, n_dims, nullptr,
                                    nullptr, cond_prev);
    rp->P.type.clear();
    for (size_t k = 0; k < nt; ++k) {
        h[k] = r.h[k];
        rho[k] = r.rho[k];
        nngb[k] = r.nngb[k];
        iters[k] = r.iters[k];
    }
}

// Per node: xpad and dmass (tree.h:273-276, :296-306); either may be NULL; dmass is 0 for a tree
// built without velocities.
void gzref_tree_pads(void* t, float* xpad, double* dmass) {
    const Tree& T = static_cast<RefTree*>(t)->T;
    for (size_t i = 0; i < T.nnodes(); ++i) {
        if (xpad) xpad[i] = T.xpad[i];
        if (dmass) dmass[i] = T.dmass.empty() ? 0.0 : T.dmass[i];
    }
}

// Reuse state per node (tree.h:266-317) and the walk node view (wn): any pointer may be NULL;
// vectors hold 3 values per node.
void gzref_tree_reuse(void* t, float* vmax, double* vcom, double* dp, double* dq, double* wcom,
                      double* wmass, double* wn_s, float* wn_soft, float* wn_len) {
    const Tree& T = static_cast<RefTree*>(t)->T;
    const size_t nn = T.nnodes();
    const bool inc = !T.dp_x.empty();
    for (size_t i = 0; i < nn; ++i) {
        if (vmax) vmax[i] = T.vmax[i];
        if (vcom) {
            vcom[3 * i] = inc ? T.vcom_x[i] : 0.0;
            vcom[3 * i + 1] = inc ? T.vcom_y[i] : 0.0;
            vcom[3 * i + 2] = inc ? T.vcom_z[i] : 0.0;
        }
        if (dp) {
            dp[3 * i] = inc ? T.dp_x[i] : 0.0;
            dp[3 * i + 1] = inc ? T.dp_y[i] : 0.0;
            dp[3 * i + 2] = inc ? T.dp_z[i] : 0.0;
        }
        if (dq) {
            dq[3 * i] = inc ? T.dq_x[i] : 0.0;
            dq[3 * i + 1] = inc ? T.dq_y[i] : 0.0;
            dq[3 * i + 2] = inc ? T.dq_z[i] : 0.0;
        }
        if (wcom) { wcom[3 * i] = T.wn[i].cx; wcom[3 * i + 1] = T.wn[i].cy; wcom[3 * i + 2] = T.wn[i].cz; }
        if (wmass) wmass[i] = T.wn[i].mass;
        if (wn_s) wn_s[i] = T.wn[i].s;
        if (wn_soft) wn_soft[i] = T.wn[i].soft;
        if (wn_len) wn_len[i] = T.wn[i].len;
    }
}

// shmem::ngb_search (hydro.h:73-75, hydro.cc:37-108) for n_q centres, non-periodic, with the tree
// aged tau (t_since_build). CSR output: start[q] .. start[q+1] in found; returns the total, and
// writes found only while it fits in cap.
size_t gzref_ngb_search(void* t, void* particles, const double* cx, const double* cy,
                        const double* cz, const double* radius, size_t n_q, double tau,
                        size_t* start, uint32_t* found, size_t cap) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    rt->T.t_since_build = tau;
    std::vector<uint32_t> all, one;
    start[0] = 0;
    for (size_t q = 0; q < n_q; ++q) {
        one.clear();
        ngb_search(rt->T, rp->P, Vec3d{cx[q], cy[q], cz[q]}, radius[q], one);
        all.insert(all.end(), one.begin(), one.end());
        start[q + 1] = all.size();
    }
    if (found && all.size() <= cap) std::memcpy(found, all.data(), all.size() * sizeof(uint32_t));
    return all.size();
}

// Tree::init_hmax (tree.cc:332-338) from h[0..nh) with the first n_gas particles gas; the node
// values are copied to hmax_out (nnodes floats) when non-NULL.
void gzref_init_hmax(void* t, const double* h, size_t nh, size_t n_gas, float* hmax_out) {
    auto* rt = static_cast<RefTree*>(t);
    rt->T.init_hmax(std::vector<double>(h, h + nh), n_gas);
    if (hmax_out)
        std::memcpy(hmax_out, rt->T.hmax.data(), rt->T.hmax.size() * sizeof(float));
}

// Tree::raise_hmax (tree.cc:293-301) for the leaf of each listed particle, to h[i]; the node
// values are copied to hmax_out when non-NULL.
void gzref_raise_hmax(void* t, const uint32_t* idx, size_t n, const double* h, float* hmax_out) {
    auto* rt = static_cast<RefTree*>(t);
    for (size_t k = 0; k < n; ++k) {
        const uint32_t i = idx[k];
        if (i < rt->T.leaf_of.size() && h[i] > 0) rt->T.raise_hmax(rt->T.leaf_of[i], h[i]);
    }
    if (hmax_out)
        std::memcpy(hmax_out, rt->T.hmax.data(), rt->T.hmax.size() * sizeof(float));
}

// shmem::ngb_search_reverse (hydro.h:81-83, hydro.cc:110-148) for n_q centres with kernel radii
// hq, non-periodic, the tree aged tau; CSR output as gzref_ngb_search.
size_t gzref_ngb_search_reverse(void* t, void* particles, const double* h, size_t n_gas,
                                const double* cx, const double* cy, const double* cz,
                                const double* hq, size_t n_q, double tau, size_t* start,
                                uint32_t* found, size_t cap) {
    auto* rt = static_cast<RefTree*>(t);
    auto* rp = static_cast<RefParticles*>(particles);
    rt->T.t_since_build = tau;
    std::vector<uint32_t> all, one;
    start[0] = 0;
    for (size_t q = 0; q < n_q; ++q) {
        one.clear();
        ngb_search_reverse(rt->T, rp->P, h, n_gas, Vec3d{cx[q], cy[q], cz[q]}, hq[q], one);
        all.insert(all.end(), one.begin(), one.end());
        start[q + 1] = all.size();
    }
    if (found && all.size() <= cap) std::memcpy(found, all.data(), all.size() * sizeof(uint32_t));
    return all.size();
}

int gzref_max_threads(void) { return omp_get_max_threads(); }
void gzref_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
