// gz_shmem.hpp -- the shmem/tree.h build and force calls in C++, over the C ABI of gz_grav.h.
//
// A program written against shmem::build / accel / accel_grouped / potential / accel_brute
// (shmem/tree.h:36-50, 371-451) switches to the device tree by including this header and writing
// gz_shmem:: for shmem::. The calls take and return host std::vectors as the reference's do:
//   * Particles is the reference's SoA; build() copies it to the device and keeps the copy;
//   * a walk reads the particles as they are at the call (the reference's walks read the live
//     arrays, tree.cc:787-788): the copy is refreshed from P and the tree's sorted source copy
//     re-read (gz_tree_update_sources) before every walk, so a caller may drift P between walks;
//   * outputs are resized to targets.size() and overwritten, accelerations include G, the tidal
//     tensor does not (tree.h:414-440), aold is per target in no-G units;
//   * with no options set, accel_grouped is the parity walk (FP64, the reference criterion, one
//     group per leaf run at batch 0); set_walk_options() selects the FP32 or learned walks;
//   * errors throw gz_shmem::Error with the library's message (the reference exits or asserts);
//     what the library does not implement (zeta corrections, lazy drift, sink-direct summation)
//     throws instead of silently differing.
//
// Header-only. Compile with the HIP runtime headers (-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include)
// and link libgzgrav.so and libamdhip64. All calls share one library context on the HIP device
// current at the first call; they are not meant to be called from several threads at once.
#ifndef GZ_SHMEM_HPP
#define GZ_SHMEM_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gz_grav.h"

namespace gz_shmem {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// shmem::Particles (tree.h:36-50). zeta must stay empty (the corrections are not implemented);
// an empty type means all gas.
struct Particles {
    std::vector<double> x, y, z, m, soft;
    std::vector<double> zeta;
    std::vector<uint8_t> type;
    size_t size() const { return m.size(); }
};

// shmem::BuildTimes (tree.h:371), in milliseconds of device time.
struct BuildTimes {
    double bbox = 0, keys = 0, sort = 0, recurse = 0, links = 0, pack = 0, total = 0;
};

// Storage of the reference's SymmetricTensor2<double> (xx yy zz xy yz xz) and Vec3<double>.
struct SymTensor3d {
    double data[6];
};
struct Vec3d {
    double data[3];
};

// The reference's lazy-drift callback type (tree.h:183-188): accepted only as nullptr.
struct LazyDrift;

namespace detail {

inline void check(gz_status s, const char* what) {
    if (s != GZ_OK) throw Error(std::string(what) + ": " + gz_last_error());
}

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DevFree {
    void operator()(void* p) const {
        if (p) (void)hipFree(p);
    }
};

// A device array that grows to the largest size asked of it.
template <class T> class DevArray {
  public:
    T* get() const { return static_cast<T*>(mem_.get()); }
    T* reserve(size_t count) {
        if (count > cap_) {
            void* p = nullptr;
            check_hip(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
            mem_.reset(p);
            cap_ = count;
        }
        return get();
    }
    T* upload(const T* host, size_t count) {
        reserve(count);
        if (count)
            check_hip(hipMemcpy(get(), host, count * sizeof(T), hipMemcpyHostToDevice),
                      "copy to device");
        return get();
    }
    T* upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
    void download(T* host, size_t count) const {
        if (count)
            check_hip(hipMemcpy(host, get(), count * sizeof(T), hipMemcpyDeviceToHost),
                      "copy to host");
    }

  private:
    std::unique_ptr<void, DevFree> mem_;
    size_t cap_ = 0;
};

struct CtxFree {
    void operator()(gz_ctx* c) const {
        if (c) (void)gz_ctx_destroy(c);
    }
};

// The one library context (its own stream) on the device current at the first call.
inline gz_ctx* ctx() {
    static std::unique_ptr<gz_ctx, CtxFree> c([] {
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        gz_ctx* x = nullptr;
        check(gz_ctx_create(dev, nullptr, &x), "gz_ctx_create");
        return x;
    }());
    return c.get();
}

inline void sync() { check(gz_ctx_synchronize(ctx()), "gz_ctx_synchronize"); }

// Device copy of a Particles.
class DevParticles {
  public:
    gz_particles upload(const Particles& P) {
        const size_t n = P.size();
        if (P.x.size() != n || P.y.size() != n || P.z.size() != n ||
            (!P.soft.empty() && P.soft.size() != n) || (!P.type.empty() && P.type.size() != n))
            throw Error("Particles: x, y, z, m (and soft, type when given) differ in length");
        if (!P.zeta.empty()) throw Error("Particles: zeta corrections are not supported");
        gz_particles p{};
        p.n = n;
        p.x = x_.upload(P.x);
        p.y = y_.upload(P.y);
        p.z = z_.upload(P.z);
        p.m = m_.upload(P.m);
        p.soft = P.soft.empty() ? nullptr : soft_.upload(P.soft);
        p.zeta = nullptr;
        p.type = P.type.empty() ? nullptr : type_.upload(P.type);
        return p;
    }

  private:
    DevArray<double> x_, y_, z_, m_, soft_;
    DevArray<uint8_t> type_;
};

// Per-call device buffers (targets, aold, outputs), reused between calls.
struct Workspace {
    DevArray<uint32_t> targets, idx;
    DevArray<double> aold, a, extra, v, vlast;
    DevParticles particles;
};
inline Workspace& ws() {
    static Workspace w;
    return w;
}

struct TreeFree {
    void operator()(gz_tree* t) const {
        if (t) (void)gz_tree_free(t);
    }
};

} // namespace detail

class Tree;
Tree build(const Particles& P, BuildTimes* bt = nullptr, const double* const* vel = nullptr,
           bool want_vcom = false, long long randomize_seed = -1);

// shmem::Tree, device-resident. Host copies of its arrays on request; the reuse calls of
// tree.h:250-331 in batched form (the reference's engine calls them once per particle).
class Tree {
  public:
    Tree() = default;
    gz_tree* handle() const { return t_.get(); }
    size_t n() const { return info().n; }
    int nnodes() const { return (int)info().nnodes; }
    int root() const { return (int)info().root; }
    double t_since_build = 0.0; // engine time since the build (tree.h:300)

    gz_tree_info info() const {
        gz_tree_info i{};
        detail::check(gz_tree_get_info(t_.get(), &i), "gz_tree_get_info");
        return i;
    }
    // orderbuf (Morton order of the live particles), rank (UINT32_MAX for tombstones) and
    // leaf_of (-1 for tombstones), as the reference keeps them (tree.h:239-267)
    std::vector<uint32_t> orderbuf() const {
        std::vector<uint32_t> o((size_t)info().n_live);
        detail::check(gz_tree_download_particles(t_.get(), o.data(), nullptr, nullptr, nullptr),
                      "gz_tree_download_particles");
        return o;
    }
    std::vector<uint32_t> rank() const {
        std::vector<uint32_t> r(n());
        detail::check(gz_tree_download_particles(t_.get(), nullptr, r.data(), nullptr, nullptr),
                      "gz_tree_download_particles");
        return r;
    }
    std::vector<int> leaf_of() const {
        std::vector<int> l(n());
        detail::check(gz_tree_download_particles(t_.get(), nullptr, nullptr, l.data(), nullptr),
                      "gz_tree_download_particles");
        return l;
    }

    // Tree::refresh_walk_nodes (tree.cc:260-274)
    void refresh_walk_nodes(double tau) {
        detail::check(gz_tree_refresh_walk_nodes(detail::ctx(), t_.get(), tau),
                      "gz_tree_refresh_walk_nodes");
        detail::sync();
    }
    // Tree::init_hmax (tree.h:310-312, tree.cc:332-338): node hmax from the gas kernel radii h
    // (one per particle; the first n_gas are gas), for ngb_search_reverse
    void init_hmax(const std::vector<double>& h, size_t n_gas) {
        auto& w = detail::ws();
        const double* d_h = w.extra.upload(h);
        detail::check(gz_tree_init_hmax(detail::ctx(), t_.get(), d_h, h.size(), n_gas),
                      "gz_tree_init_hmax");
    }
    // Tree::raise_hmax (tree.cc:293-301) for the leaf of each listed particle, to h[i]
    void raise_hmax(const std::vector<uint32_t>& particles, const std::vector<double>& h) {
        if (h.size() != n()) throw Error("raise_hmax: h must hold one value per particle");
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        const double* d_h = w.extra.upload(h);
        detail::check(gz_tree_raise_hmax(detail::ctx(), t_.get(), d_idx, particles.size(), d_h),
                      "gz_tree_raise_hmax");
    }
    // Tree::raise_soft (tree.cc:311-319) for the leaf of each listed particle, to soft[i]
    void raise_soft(const std::vector<uint32_t>& particles, const std::vector<double>& soft) {
        if (soft.size() != n()) throw Error("raise_soft: soft must hold one value per particle");
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        const double* d_soft = w.aold.upload(soft);
        detail::check(gz_tree_raise_soft(detail::ctx(), t_.get(), d_idx, particles.size(), d_soft),
                      "gz_tree_raise_soft");
        detail::sync();
    }
    // The kick feed of mfm.cc:3798-3811 (Tree::raise_vmax and Tree::kick_node) for the active
    // particles: vel = {vx, vy, vz} after the kick, vlast = {vx, vy, vz} at the previous feed
    // (updated to vel), m the masses, tau = t_since_build. vlast may be nullptr (speed bound
    // only, as for a tree built without velocities).
    void kick_nodes(const std::vector<uint32_t>& active, const double* const* vel,
                    const std::vector<double>& m, double* const* vlast, double tau) {
        const size_t N = n(), na = active.size();
        if (m.size() != N) throw Error("kick_nodes: m must hold one value per particle");
        auto& w = detail::ws();
        const uint32_t* d_act = w.idx.upload(active);
        double* v = w.v.reserve(3 * N);
        double* vl = vlast ? w.vlast.reserve(3 * N) : nullptr;
        for (int d = 0; d < 3; ++d) {
            detail::check_hip(hipMemcpy(v + d * N, vel[d], N * sizeof(double),
                                        hipMemcpyHostToDevice), "copy to device");
            if (vl)
                detail::check_hip(hipMemcpy(vl + d * N, vlast[d], N * sizeof(double),
                                            hipMemcpyHostToDevice), "copy to device");
        }
        const double* d_m = w.aold.upload(m);
        detail::check(gz_tree_kick_nodes(detail::ctx(), t_.get(), d_act, na, v, v + N, v + 2 * N,
                                         d_m, vl, vl ? vl + N : nullptr,
                                         vl ? vl + 2 * N : nullptr, tau),
                      "gz_tree_kick_nodes");
        detail::sync();
        if (vl)
            for (int d = 0; d < 3; ++d)
                detail::check_hip(hipMemcpy(vlast[d], vl + d * N, N * sizeof(double),
                                            hipMemcpyDeviceToHost), "copy to host");
    }

    // Tree::mass_event (tree.h:321-323, tree.cc:239-251) for each event k in order: a particle
    // of mass dm[k] (negative removes one) at x[k] moving at v[k], appearing at tau, in the
    // build-time leaf of particle particles[k] (where the reference takes leaf_of[i]); x is
    // unwrapped against each ancestor's centre of mass in a box of side box (0: open).
    void mass_events(const std::vector<uint32_t>& particles, const std::vector<double>& dm,
                     const std::vector<Vec3d>& x, const std::vector<Vec3d>& v, double tau,
                     double box = 0.0) {
        const size_t ne = particles.size();
        if (dm.size() != ne || x.size() != ne || v.size() != ne)
            throw Error("mass_events: one mass, position and velocity per event");
        if (ne == 0) return;
        std::vector<double> buf(7 * ne);
        for (size_t k = 0; k < ne; ++k) {
            buf[k] = dm[k];
            for (int d = 0; d < 3; ++d) {
                buf[(1 + d) * ne + k] = x[k].data[d];
                buf[(4 + d) * ne + k] = v[k].data[d];
            }
        }
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        const double* b = w.extra.upload(buf);
        detail::check(gz_tree_mass_events(detail::ctx(), t_.get(), d_idx, ne, b, b + ne,
                                          b + 2 * ne, b + 3 * ne, b + 4 * ne, b + 5 * ne,
                                          b + 6 * ne, tau, box),
                      "gz_tree_mass_events");
    }
    // Tree::shift_node (tree.h:324-325, tree.cc:253-258) for each event k: a particle of mass
    // m[k] displaced by dx[k] without a change of velocity.
    void shift_nodes(const std::vector<uint32_t>& particles, const std::vector<double>& m,
                     const std::vector<Vec3d>& dx) {
        const size_t ne = particles.size();
        if (m.size() != ne || dx.size() != ne)
            throw Error("shift_nodes: one mass and displacement per event");
        if (ne == 0) return;
        std::vector<double> buf(4 * ne);
        for (size_t k = 0; k < ne; ++k) {
            buf[k] = m[k];
            for (int d = 0; d < 3; ++d) buf[(1 + d) * ne + k] = dx[k].data[d];
        }
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        const double* b = w.extra.upload(buf);
        detail::check(gz_tree_shift_nodes(detail::ctx(), t_.get(), d_idx, ne, b, b + ne,
                                          b + 2 * ne, b + 3 * ne),
                      "gz_tree_shift_nodes");
    }
    // Tree::raise_pad (tree.h:316-317, tree.cc:323-330) for each event k: a displacement d[k].
    void raise_pad(const std::vector<uint32_t>& particles, const std::vector<double>& d) {
        if (d.size() != particles.size()) throw Error("raise_pad: one displacement per event");
        if (d.empty()) return;
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        const double* d_d = w.extra.upload(d);
        detail::check(gz_tree_raise_pad(detail::ctx(), t_.get(), d_idx, d.size(), d_d),
                      "gz_tree_raise_pad");
    }
    // Tree::raise_vmax (tree.h:303-309, tree.cc:276-285) for each event k: speed[k] >= 0.
    void raise_vmax(const std::vector<uint32_t>& particles, const std::vector<float>& speed) {
        if (speed.size() != particles.size()) throw Error("raise_vmax: one speed per event");
        if (speed.empty()) return;
        auto& w = detail::ws();
        const uint32_t* d_idx = w.idx.upload(particles);
        detail::DevArray<float> d_s;
        const float* ds = d_s.upload(speed);
        detail::check(gz_tree_raise_vmax(detail::ctx(), t_.get(), d_idx, speed.size(), ds),
                      "gz_tree_raise_vmax");
        detail::sync();
    }

  private:
    friend Tree build(const Particles&, BuildTimes*, const double* const*, bool, long long);
    friend struct WalkAccess;
    std::shared_ptr<gz_tree> t_;
};

// Walk options for the calls below; nullptr (the default) is the parity walk.
inline const gz_walk_opts*& walk_options() {
    static const gz_walk_opts* o = nullptr;
    return o;
}
inline void set_walk_options(const gz_walk_opts* o) { walk_options() = o; }

// shmem::build (tree.h:384-385, tree.cc:350-526). want_vcom is implied by vel: a tree built with
// velocities keeps the speed bound and the centre-of-mass velocities both.
inline Tree build(const Particles& P, BuildTimes* bt, const double* const* vel, bool want_vcom,
                  long long randomize_seed) {
    (void)want_vcom;
    auto& w = detail::ws();
    const gz_particles p = w.particles.upload(P);
    gz_build_opts o{};
    o.randomize_seed = randomize_seed;
    const size_t N = P.size();
    if (vel) {
        double* v = w.v.reserve(3 * N);
        for (int d = 0; d < 3; ++d)
            detail::check_hip(hipMemcpy(v + d * N, vel[d], N * sizeof(double),
                                        hipMemcpyHostToDevice), "copy to device");
        o.vx = v;
        o.vy = v + N;
        o.vz = v + 2 * N;
    }
    gz_tree* t = nullptr;
    detail::check(gz_tree_build(detail::ctx(), &p, &o, &t), "gz_tree_build");
    Tree T;
    T.t_ = std::shared_ptr<gz_tree>(t, detail::TreeFree());
    if (bt) {
        const gz_build_times& s = T.info().times;
        *bt = BuildTimes{s.bbox, s.keys, s.sort, s.recurse, s.links, s.pack, s.total};
    }
    return T;
}

struct WalkAccess {
    // Particles to the device and into the tree's sorted source copy; targets and aold staged.
    static gz_particles stage(const Tree& T, const Particles& P, const char* who) {
        if (!T.t_) throw Error(std::string(who) + ": empty tree");
        if (P.size() != T.n())
            throw Error(std::string(who) + ": the particles differ in number from the tree's");
        const gz_particles p = detail::ws().particles.upload(P);
        detail::check(gz_tree_update_sources(detail::ctx(), T.t_.get(), &p),
                      "gz_tree_update_sources");
        return p;
    }
    static gz_tree* tree(const Tree& T) { return T.t_.get(); }
};

namespace detail {
inline const uint32_t* stage_targets(const std::vector<uint32_t>& targets) {
    return ws().targets.upload(targets);
}
inline const double* stage_aold(const double* aold, size_t nt) {
    return aold ? ws().aold.upload(aold, nt) : nullptr;
}
inline void fetch3(double* a, size_t nt, std::vector<double>& ax, std::vector<double>& ay,
                   std::vector<double>& az) {
    ax.resize(nt);
    ay.resize(nt);
    az.resize(nt);
    sync();
    check_hip(hipMemcpy(ax.data(), a, nt * sizeof(double), hipMemcpyDeviceToHost), "copy");
    check_hip(hipMemcpy(ay.data(), a + nt, nt * sizeof(double), hipMemcpyDeviceToHost), "copy");
    check_hip(hipMemcpy(az.data(), a + 2 * nt, nt * sizeof(double), hipMemcpyDeviceToHost),
              "copy");
}
} // namespace detail

// shmem::accel (tree.h:392-394): the per-target walk, leaves always summed pair by pair.
inline void accel(const Tree& T, const Particles& P, const std::vector<uint32_t>& targets,
                  double theta, double G, std::vector<double>& ax, std::vector<double>& ay,
                  std::vector<double>& az, const double* aold = nullptr) {
    const size_t nt = targets.size();
    const gz_particles p = WalkAccess::stage(T, P, "accel");
    const uint32_t* d_t = detail::stage_targets(targets);
    const double* d_ao = detail::stage_aold(aold, nt);
    double* a = detail::ws().a.reserve(3 * nt + 1);
    detail::check(gz_tree_accel(detail::ctx(), WalkAccess::tree(T), &p, d_t, nt, theta, G, d_ao,
                                a, a + nt, a + 2 * nt),
                  "gz_tree_accel");
    detail::fetch3(a, nt, ax, ay, az);
}

// shmem::accel_grouped (tree.h:428-440). tidal and jerk come from the FP64 walk with the
// reference criterion; jerk needs vel and a tree built with velocities.
inline void accel_grouped(const Tree& T, const Particles& P, const std::vector<uint32_t>& targets,
                          double theta, double G, int batch, std::vector<double>& ax,
                          std::vector<double>& ay, std::vector<double>& az,
                          std::vector<SymTensor3d>* tidal = nullptr, const double* aold = nullptr,
                          const LazyDrift* lazy = nullptr, std::vector<Vec3d>* jerk = nullptr,
                          const double* const* vel = nullptr, double sink_direct_radius = 0.0) {
    if (lazy) throw Error("accel_grouped: lazy drift is not supported");
    if (sink_direct_radius != 0.0)
        throw Error("accel_grouped: sink-direct summation is not supported");
    if (jerk && !vel) throw Error("accel_grouped: the jerk needs the velocities (vel)");
    const size_t nt = targets.size();
    const gz_particles p = WalkAccess::stage(T, P, "accel_grouped");
    const uint32_t* d_t = detail::stage_targets(targets);
    const double* d_ao = detail::stage_aold(aold, nt);
    auto& w = detail::ws();
    double* a = w.a.reserve(3 * nt + 1);
    if (tidal || jerk) {
        double* ex = w.extra.reserve(9 * nt + 1);
        const double *vx = nullptr, *vy = nullptr, *vz = nullptr;
        if (vel) {
            const size_t N = P.size();
            double* v = w.v.reserve(3 * N);
            for (int d = 0; d < 3; ++d)
                detail::check_hip(hipMemcpy(v + d * N, vel[d], N * sizeof(double),
                                            hipMemcpyHostToDevice), "copy to device");
            vx = v;
            vy = v + N;
            vz = v + 2 * N;
        }
        detail::check(gz_tree_accel_grouped_tidal(detail::ctx(), WalkAccess::tree(T), &p, d_t,
                                                  nt, theta, G, batch, d_ao, vx, vy, vz, a,
                                                  a + nt, a + 2 * nt, ex,
                                                  jerk ? ex + 6 * nt : nullptr),
                      "gz_tree_accel_grouped_tidal");
        detail::fetch3(a, nt, ax, ay, az);
        std::vector<double> h(9 * nt);
        w.extra.download(h.data(), (jerk ? 9 : 6) * nt);
        if (tidal) {
            tidal->resize(nt);
            for (size_t k = 0; k < nt; ++k)
                for (int c = 0; c < 6; ++c) (*tidal)[k].data[c] = h[c * nt + k];
        }
        if (jerk) {
            jerk->resize(nt);
            for (size_t k = 0; k < nt; ++k)
                for (int c = 0; c < 3; ++c) (*jerk)[k].data[c] = h[(6 + c) * nt + k];
        }
        return;
    }
    detail::check(gz_tree_accel_grouped(detail::ctx(), WalkAccess::tree(T), &p, d_t, nt, theta,
                                        G, batch, d_ao, a, a + nt, a + 2 * nt, walk_options(),
                                        nullptr),
                  "gz_tree_accel_grouped");
    detail::fetch3(a, nt, ax, ay, az);
}

// shmem::potential (tree.h:446-447), G included.
inline void potential(const Tree& T, const Particles& P, const std::vector<uint32_t>& targets,
                      double theta, double G, std::vector<double>& phi,
                      const double* aold = nullptr) {
    const size_t nt = targets.size();
    const gz_particles p = WalkAccess::stage(T, P, "potential");
    const uint32_t* d_t = detail::stage_targets(targets);
    const double* d_ao = detail::stage_aold(aold, nt);
    double* d_phi = detail::ws().a.reserve(nt + 1);
    detail::check(gz_tree_potential(detail::ctx(), WalkAccess::tree(T), &p, d_t, nt, theta, G,
                                    d_ao, d_phi),
                  "gz_tree_potential");
    phi.resize(nt);
    detail::sync();
    detail::ws().a.download(phi.data(), nt);
}

// shmem::ngb_search (hydro.h:73-75, hydro.cc:37-108): APPENDS to found the particles within
// radius of centre, in the tree's Morton order, with node boxes widened by vmax * t_since_build.
// Periodic boxes (box != 0) and lazy drift are not supported. One launch per call: callers with
// many centres use ngb_search_many.
inline void ngb_search_many(const Tree& T, const Particles& P, const std::vector<Vec3d>& centres,
                            const std::vector<double>& radius,
                            std::vector<std::vector<uint32_t>>& found) {
    const size_t nq = centres.size();
    if (radius.size() != nq) throw Error("ngb_search: one radius per centre");
    found.resize(nq);
    if (nq == 0) return;
    const gz_particles p = WalkAccess::stage(T, P, "ngb_search");
    (void)p;
    std::vector<double> c(3 * nq);
    for (size_t q = 0; q < nq; ++q)
        for (int d = 0; d < 3; ++d) c[d * nq + q] = centres[q].data[d];
    auto& w = detail::ws();
    double* d_c = w.v.upload(c);
    const double* d_r = w.aold.upload(radius);
    detail::DevArray<size_t> start;
    start.reserve(nq + 1);
    // the lists into a buffer of 32 entries per query, searched again only if they need more
    size_t total = 0, cap = std::max<size_t>(32 * nq, 1);
    detail::DevArray<uint32_t> hits;
    for (;;) {
        hits.reserve(cap);
        detail::check(gz_tree_ngb_search(detail::ctx(), WalkAccess::tree(T), d_c, d_c + nq,
                                         d_c + 2 * nq, d_r, nq, T.t_since_build, start.get(),
                                         hits.get(), cap, &total),
                      "gz_tree_ngb_search");
        if (total <= cap) break;
        cap = total;
    }
    detail::sync();
    std::vector<size_t> hs(nq + 1);
    std::vector<uint32_t> hh(total);
    start.download(hs.data(), nq + 1);
    hits.download(hh.data(), total);
    for (size_t q = 0; q < nq; ++q)
        found[q].insert(found[q].end(), hh.begin() + (std::ptrdiff_t)hs[q],
                        hh.begin() + (std::ptrdiff_t)hs[q + 1]);
}

inline void ngb_search(const Tree& T, const Particles& P, const Vec3d& centre, double radius,
                       std::vector<uint32_t>& found, double box = 0.0,
                       const LazyDrift* lazy = nullptr) {
    if (box != 0.0) throw Error("ngb_search: periodic boxes are not supported");
    if (lazy) throw Error("ngb_search: lazy drift is not supported");
    std::vector<std::vector<uint32_t>> f(1);
    ngb_search_many(T, P, std::vector<Vec3d>{centre}, std::vector<double>{radius}, f);
    found.insert(found.end(), f[0].begin(), f[0].end());
}

// shmem::ngb_search_reverse (hydro.h:78-83, hydro.cc:110-148): APPENDS to found[q] the gas
// particles j (j < n_gas) with h_q[q] <= r < h[j], whose own kernel reaches centre q, in Morton
// order; h holds at least n_gas radii by particle; needs Tree::init_hmax. Node boxes are
// widened by vmax * t_since_build. One launch for all centres.
inline void ngb_search_reverse_many(const Tree& T, const Particles& P, const double* h,
                                    size_t n_gas, const std::vector<Vec3d>& centres,
                                    const std::vector<double>& h_q,
                                    std::vector<std::vector<uint32_t>>& found) {
    const size_t nq = centres.size();
    if (h_q.size() != nq) throw Error("ngb_search_reverse: one h_q per centre");
    found.resize(nq);
    if (nq == 0) return;
    const gz_particles p = WalkAccess::stage(T, P, "ngb_search_reverse");
    (void)p;
    std::vector<double> c(3 * nq);
    for (size_t q = 0; q < nq; ++q)
        for (int d = 0; d < 3; ++d) c[d * nq + q] = centres[q].data[d];
    auto& w = detail::ws();
    double* d_c = w.v.upload(c);
    const double* d_r = w.aold.upload(h_q);
    const double* d_h = w.extra.upload(std::vector<double>(h, h + n_gas));
    detail::DevArray<size_t> start;
    start.reserve(nq + 1);
    // the lists into a buffer of 32 entries per query, searched again only if they need more
    size_t total = 0, cap = std::max<size_t>(32 * nq, 1);
    detail::DevArray<uint32_t> hits;
    for (;;) {
        hits.reserve(cap);
        detail::check(gz_tree_ngb_search_reverse(detail::ctx(), WalkAccess::tree(T), d_h, n_gas,
                                                 d_c, d_c + nq, d_c + 2 * nq, d_r, nq,
                                                 T.t_since_build, start.get(), hits.get(), cap,
                                                 &total),
                      "gz_tree_ngb_search_reverse");
        if (total <= cap) break;
        cap = total;
    }
    detail::sync();
    std::vector<size_t> hs(nq + 1);
    std::vector<uint32_t> hh(total);
    start.download(hs.data(), nq + 1);
    hits.download(hh.data(), total);
    for (size_t q = 0; q < nq; ++q)
        found[q].insert(found[q].end(), hh.begin() + (std::ptrdiff_t)hs[q],
                        hh.begin() + (std::ptrdiff_t)hs[q + 1]);
}

inline void ngb_search_reverse(const Tree& T, const Particles& P, const double* h, size_t n_gas,
                               const Vec3d& centre, double h_i, std::vector<uint32_t>& found,
                               double box = 0.0, const LazyDrift* lazy = nullptr) {
    if (box != 0.0) throw Error("ngb_search_reverse: periodic boxes are not supported");
    if (lazy) throw Error("ngb_search_reverse: lazy drift is not supported");
    std::vector<std::vector<uint32_t>> f(1);
    ngb_search_reverse_many(T, P, h, n_gas, std::vector<Vec3d>{centre}, std::vector<double>{h_i},
                            f);
    found.insert(found.end(), f[0].begin(), f[0].end());
}

// shmem::DensityResult and shmem::NeighborCache (hydro.h:86-108).
struct DensityResult {
    std::vector<double> h, rho;
    std::vector<int> nngb, iters;
};
struct NeighborCache {
    std::vector<uint32_t> flat;
    std::vector<size_t> start;
    bool valid = false;
    void clear() {
        flat.clear();
        start.clear();
        valid = false;
    }
};

// shmem::density (hydro.h:115-119, hydro.cc:150-327): h and rho for every target, with the
// gas neighbours of Particles::type. Periodic boxes and lazy drift are not supported. With a
// cache, each target's neighbour list at its final h (the reference's last search) is filled by
// one more search.
inline DensityResult density(const Tree& T, const Particles& P,
                             const std::vector<uint32_t>& targets, double des_ngb, double ngb_tol,
                             const std::vector<double>& h_start, double box = 0.0, int n_dims = 3,
                             NeighborCache* cache = nullptr, const LazyDrift* lazy = nullptr,
                             const double* cond_prev = nullptr) {
    if (box != 0.0) throw Error("density: periodic boxes are not supported");
    if (lazy) throw Error("density: lazy drift is not supported");
    const size_t nt = targets.size();
    if (!h_start.empty() && h_start.size() != nt) throw Error("density: one h_start per target");
    DensityResult r;
    r.h.resize(nt);
    r.rho.resize(nt);
    r.nngb.resize(nt);
    r.iters.resize(nt);
    if (cache) cache->clear();
    if (nt > 0) {
        const gz_particles p = WalkAccess::stage(T, P, "density");
        auto& w = detail::ws();
        const uint32_t* d_t = detail::stage_targets(targets);
        const double* d_hs = h_start.empty() ? nullptr : w.aold.upload(h_start);
        detail::DevArray<double> d_cond, d_out;
        const double* d_c = cond_prev ? d_cond.upload(cond_prev, P.size()) : nullptr;
        double* o = d_out.reserve(3 * nt);
        int* oi = reinterpret_cast<int*>(o + 2 * nt);
        detail::check(gz_tree_density(detail::ctx(), WalkAccess::tree(T), &p, d_t, nt, des_ngb,
                                      ngb_tol, d_hs, n_dims, d_c, T.t_since_build, o, o + nt, oi,
                                      oi + nt),
                      "gz_tree_density");
        detail::sync();
        d_out.download(r.h.data(), nt);
        detail::check_hip(hipMemcpy(r.rho.data(), o + nt, nt * sizeof(double),
                                    hipMemcpyDeviceToHost), "copy");
        detail::check_hip(hipMemcpy(r.nngb.data(), oi, nt * sizeof(int), hipMemcpyDeviceToHost),
                          "copy");
        detail::check_hip(hipMemcpy(r.iters.data(), oi + nt, nt * sizeof(int),
                                    hipMemcpyDeviceToHost), "copy");
    }
    if (cache) {
        std::vector<Vec3d> c(nt);
        for (size_t k = 0; k < nt; ++k) c[k] = Vec3d{{P.x[targets[k]], P.y[targets[k]], P.z[targets[k]]}};
        std::vector<std::vector<uint32_t>> f;
        ngb_search_many(T, P, c, r.h, f);
        cache->start.assign(nt + 1, 0);
        for (size_t k = 0; k < nt; ++k) {
            cache->start[k + 1] = cache->start[k] + f[k].size();
            cache->flat.insert(cache->flat.end(), f[k].begin(), f[k].end());
        }
        cache->valid = true;
    }
    return r;
}

// shmem::accel_brute (tree.h:450-451): the exact FP64 sum over all particles, G included.
inline void accel_brute(const Particles& P, const std::vector<uint32_t>& targets, double G,
                        std::vector<double>& ax, std::vector<double>& ay,
                        std::vector<double>& az) {
    const size_t nt = targets.size();
    const gz_particles p = detail::ws().particles.upload(P);
    const uint32_t* d_t = detail::stage_targets(targets);
    double* a = detail::ws().a.reserve(3 * nt + 1);
    detail::check(gz_direct_sum(detail::ctx(), &p, d_t, nt, G, a, a + nt, a + 2 * nt),
                  "gz_direct_sum");
    detail::fetch3(a, nt, ax, ay, az);
}

} // namespace gz_shmem

#endif // GZ_SHMEM_HPP
