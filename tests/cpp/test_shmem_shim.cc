// The shmem/tree.h calls of include/gz_shmem.hpp on the device, checked against the oracle's C
// restatement (oracle/gz_oracle.h, itself pinned bit for bit to the reference's shmem/tree.cc) on
// the same Plummer sphere, in the manner of the reference's own shmem/test_tree.cc.
//
// Tolerances: orderbuf and node count exact; the FP64 walks 1e-12 relative (the device sums
// interactions in another order); the direct sum 1e-13; the tidal tensor 1e-11 of its largest
// component; the FP32 walk 1e-4 of the FP64 walk. Exit status 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "gz_oracle.h"
#include "gz_shmem.hpp"

namespace {

int g_failures = 0;

void expect(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    if (!ok) ++g_failures;
}

// Plummer sphere of scale radius a: mass fraction u in [0, 0.999), r = a (u^-2/3 - 1)^-1/2,
// isotropic directions; equal masses summing to 1.
gz_shmem::Particles plummer(size_t n, double a, double eps, unsigned seed) {
    gz_shmem::Particles P;
    P.x.resize(n);
    P.y.resize(n);
    P.z.resize(n);
    P.m.assign(n, 1.0 / (double)n);
    P.soft.assign(n, eps);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (size_t i = 0; i < n; ++i) {
        const double u = U(rng) * 0.999;
        const double r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        const double ct = 2.0 * U(rng) - 1.0, st = std::sqrt(1.0 - ct * ct);
        const double ph = 2.0 * M_PI * U(rng);
        P.x[i] = r * st * std::cos(ph);
        P.y[i] = r * st * std::sin(ph);
        P.z[i] = r * ct;
    }
    return P;
}

gzo_particles view(const gz_shmem::Particles& P) {
    return gzo_particles{P.size(), P.x.data(), P.y.data(), P.z.data(), P.m.data(),
                         P.soft.data()};
}

// Largest |a - b| / |b| over the targets.
double max_rel(const std::vector<double>& ax, const std::vector<double>& ay,
               const std::vector<double>& az, const std::vector<double>& bx,
               const std::vector<double>& by, const std::vector<double>& bz) {
    double worst = 0.0;
    for (size_t k = 0; k < bx.size(); ++k) {
        const double dx = ax[k] - bx[k], dy = ay[k] - by[k], dz = az[k] - bz[k];
        const double nb = std::sqrt(bx[k] * bx[k] + by[k] * by[k] + bz[k] * bz[k]);
        worst = std::fmax(worst, std::sqrt(dx * dx + dy * dy + dz * dz) / std::fmax(nb, 1e-300));
    }
    return worst;
}

template <class F> bool throws(F&& f) {
    try {
        f();
    } catch (const gz_shmem::Error&) {
        return true;
    }
    return false;
}

std::string fmt(const char* what, double v) {
    char b[160];
    std::snprintf(b, sizeof b, "%s (%.2e)", what, v);
    return b;
}

} // namespace

int main() {
    const double G = 1.0;
    const size_t n = 20000;
    gz_shmem::Particles P = plummer(n, 1.0, 0.01, 12345u);
    const gzo_particles hp = view(P);

    gz_shmem::BuildTimes bt;
    gz_shmem::Tree T = gz_shmem::build(P, &bt);
    void* to = gzo_build(&hp, -1);
    long long sizes[3];
    gzo_tree_sizes(to, sizes);
    expect(T.nnodes() == (int)sizes[0], "node count equals the oracle's");
    expect(bt.total > 0.0, "build times reported");

    // the target list of an all-active step: the tree's Morton order
    std::vector<uint32_t> tg = T.orderbuf();
    std::vector<uint32_t> ob((size_t)sizes[2]);
    gzo_tree_particles(to, ob.data(), nullptr, nullptr);
    expect(tg == ob, "orderbuf equals the oracle's");
    const size_t nt = tg.size();

    // accel_grouped, batch 0: the Barnes-Hut bootstrap, then the relative criterion
    std::vector<double> ax, ay, az, ox(nt), oy(nt), oz(nt);
    gz_shmem::accel_grouped(T, P, tg, 0.7, G, 0, ax, ay, az);
    gzo_accel_grouped(to, &hp, tg.data(), nt, 0.7, G, 0, nullptr, ox.data(), oy.data(),
                      oz.data(), nullptr);
    expect(ax.size() == nt, "outputs resized to the target count");
    double e = max_rel(ax, ay, az, ox, oy, oz);
    expect(e < 1e-12, fmt("accel_grouped theta=0.7 equals the oracle", e));
    std::vector<double> aold(nt);
    for (size_t k = 0; k < nt; ++k)
        aold[k] = 0.0025 * std::sqrt(ox[k] * ox[k] + oy[k] * oy[k] + oz[k] * oz[k]) / G;
    gz_shmem::accel_grouped(T, P, tg, 0.0, G, 0, ax, ay, az, nullptr, aold.data());
    gzo_accel_grouped(to, &hp, tg.data(), nt, 0.0, G, 0, aold.data(), ox.data(), oy.data(),
                      oz.data(), nullptr);
    e = max_rel(ax, ay, az, ox, oy, oz);
    expect(e < 1e-12, fmt("accel_grouped relative criterion equals the oracle", e));

    // a random subset of targets in another order, batches of 8
    std::vector<uint32_t> sub;
    for (size_t i = 0; i < n; i += 7) sub.push_back((uint32_t)((i * 7919u) % n));
    std::vector<double> sx(sub.size()), sy(sub.size()), sz(sub.size());
    gz_shmem::accel_grouped(T, P, sub, 0.5, G, 8, ax, ay, az);
    gzo_accel_grouped(to, &hp, sub.data(), sub.size(), 0.5, G, 8, nullptr, sx.data(), sy.data(),
                      sz.data(), nullptr);
    e = max_rel(ax, ay, az, sx, sy, sz);
    expect(e < 1e-12, fmt("accel_grouped batch 8 on a shuffled subset equals the oracle", e));

    // per-target walk and potential
    gz_shmem::accel(T, P, sub, 0.6, G, ax, ay, az);
    gzo_accel(to, &hp, sub.data(), sub.size(), 0.6, G, nullptr, sx.data(), sy.data(), sz.data());
    e = max_rel(ax, ay, az, sx, sy, sz);
    expect(e < 1e-12, fmt("accel equals the oracle", e));
    std::vector<double> phi, ophi(sub.size());
    gz_shmem::potential(T, P, sub, 0.6, G, phi);
    gzo_potential(to, &hp, sub.data(), sub.size(), 0.6, G, nullptr, ophi.data());
    double ep = 0.0;
    for (size_t k = 0; k < sub.size(); ++k)
        ep = std::fmax(ep, std::fabs(phi[k] - ophi[k]) / std::fabs(ophi[k]));
    expect(ep < 1e-12, fmt("potential equals the oracle", ep));

    // the judge
    std::vector<uint32_t> probe(sub.begin(), sub.begin() + 512);
    std::vector<double> bx(512), by(512), bz(512);
    gz_shmem::accel_brute(P, probe, G, ax, ay, az);
    gzo_accel_brute(&hp, probe.data(), probe.size(), G, bx.data(), by.data(), bz.data());
    e = max_rel(ax, ay, az, bx, by, bz);
    expect(e < 1e-13, fmt("accel_brute equals the oracle", e));

    // tidal tensor from the same walk
    std::vector<gz_shmem::SymTensor3d> tid;
    std::vector<double> otid(6 * sub.size());
    gz_shmem::accel_grouped(T, P, sub, 0.7, G, 0, ax, ay, az, &tid);
    gzo_accel_grouped_tj(to, &hp, sub.data(), sub.size(), 0.7, G, 0, nullptr, nullptr, nullptr,
                         nullptr, sx.data(), sy.data(), sz.data(), otid.data(), nullptr, nullptr);
    double tmax = 0.0, terr = 0.0;
    for (size_t k = 0; k < sub.size(); ++k)
        for (int c = 0; c < 6; ++c) {
            tmax = std::fmax(tmax, std::fabs(otid[c * sub.size() + k]));
            terr = std::fmax(terr, std::fabs(tid[k].data[c] - otid[c * sub.size() + k]));
        }
    expect(tid.size() == sub.size() && terr <= 1e-11 * tmax,
           fmt("tidal tensor equals the oracle", terr / tmax));
    e = max_rel(ax, ay, az, sx, sy, sz);
    expect(e < 1e-12, fmt("accelerations of the tidal walk equal the oracle", e));

    // the particles drift after the build: the walk reads them as they are now
    for (size_t i = 0; i < n; ++i) {
        P.x[i] += 1e-4 * std::sin(3.0 * (double)i);
        P.y[i] += 1e-4 * std::cos(5.0 * (double)i);
    }
    const gzo_particles hp2 = view(P);
    gz_shmem::accel_grouped(T, P, tg, 0.7, G, 0, ax, ay, az);
    gzo_accel_grouped(to, &hp2, tg.data(), nt, 0.7, G, 0, nullptr, ox.data(), oy.data(),
                      oz.data(), nullptr);
    e = max_rel(ax, ay, az, ox, oy, oz);
    expect(e < 1e-12, fmt("walk on drifted particles equals the oracle's", e));

    // the FP32 group-relative walk through the same call
    gz_walk_opts wo{GZ_PREC_FP32_REL, GZ_CRIT_REFERENCE, 32, GZ_LAUNCH_QUEUE, 0, 0, nullptr, 0};
    gz_shmem::set_walk_options(&wo);
    gz_shmem::accel_grouped(T, P, tg, 0.0, G, 0, ax, ay, az, nullptr, aold.data());
    wo.precision = GZ_PREC_FP64;
    std::vector<double> fx, fy, fz;
    gz_shmem::accel_grouped(T, P, tg, 0.0, G, 0, fx, fy, fz, nullptr, aold.data());
    gz_shmem::set_walk_options(nullptr);
    e = max_rel(ax, ay, az, fx, fy, fz);
    expect(e < 1e-4, fmt("FP32 walk within 1e-4 of the FP64 walk (32-target groups)", e));

    // errors are exceptions, never a stray read
    gz_shmem::Particles Z = P;
    Z.zeta.assign(n, 0.0);
    expect(throws([&] { gz_shmem::accel_grouped(T, Z, tg, 0.7, G, 0, ax, ay, az); }),
           "zeta corrections are refused");
    std::vector<uint32_t> bad = {0u, (uint32_t)n};
    expect(throws([&] { gz_shmem::accel_grouped(T, P, bad, 0.7, G, 0, ax, ay, az); }),
           "a target index outside the particles is refused");
    expect(throws([&] {
               gz_shmem::accel_grouped(T, P, tg, 0.7, G, 0, ax, ay, az, nullptr, nullptr,
                                       nullptr, nullptr, nullptr, 1.0);
           }),
           "sink-direct summation is refused");
    gz_shmem::Particles small = plummer(100, 1.0, 0.01, 1u);
    expect(throws([&] { gz_shmem::accel_grouped(T, small, tg, 0.7, G, 0, ax, ay, az); }),
           "particles of another size are refused");

    gzo_tree_free(to);

    // ---- reuse between rebuilds: build with velocities, one kick feed, refresh, drift, walk;
    // raise_soft; the neighbour search (momentum sums are FP64 atomics on the device: 1e-10)
    {
        const size_t m2 = 10000;
        gz_shmem::Particles Q = plummer(m2, 1.0, 0.01, 777u);
        std::vector<double> vx(m2), vy(m2), vz(m2);
        std::mt19937_64 rng(99u);
        std::normal_distribution<double> N(0.0, 0.3);
        for (size_t i = 0; i < m2; ++i) {
            vx[i] = N(rng);
            vy[i] = N(rng);
            vz[i] = N(rng);
        }
        const double* vel[3] = {vx.data(), vy.data(), vz.data()};
        gz_shmem::Tree R = gz_shmem::build(Q, nullptr, vel, true);
        const gzo_particles hq = view(Q);
        void* tq = gzo_build_vel(&hq, -1, vx.data(), vy.data(), vz.data());
        // a third of the particles are kicked at tau = 0.05
        std::vector<uint32_t> act;
        for (size_t i = 0; i < m2; i += 3) act.push_back((uint32_t)i);
        std::vector<double> lx = vx, ly = vy, lz = vz, ox2 = vx, oy2 = vy, oz2 = vz;
        std::vector<double> kx = vx, ky = vy, kz = vz;
        for (uint32_t i : act) {
            kx[i] += 0.05 * N(rng);
            ky[i] += 0.05 * N(rng);
            kz[i] += 0.05 * N(rng);
        }
        const double* kv[3] = {kx.data(), ky.data(), kz.data()};
        double* last[3] = {lx.data(), ly.data(), lz.data()};
        R.kick_nodes(act, kv, Q.m, last, 0.05);
        gzo_kick_nodes(tq, act.data(), act.size(), kx.data(), ky.data(), kz.data(), Q.m.data(),
                       ox2.data(), oy2.data(), oz2.data(), 0.05);
        expect(lx == ox2 && ly == oy2 && lz == oz2, "kick feed updates v_last as the oracle does");
        R.refresh_walk_nodes(0.1);
        gzo_refresh_walk_nodes(tq, 0.1);
        R.t_since_build = 0.1;
        // particles at tau = 0.1: v before the kick until 0.05, the kicked v after
        for (size_t i = 0; i < m2; ++i) {
            Q.x[i] += 0.05 * vx[i] + 0.05 * kx[i];
            Q.y[i] += 0.05 * vy[i] + 0.05 * ky[i];
            Q.z[i] += 0.05 * vz[i] + 0.05 * kz[i];
        }
        const gzo_particles hq2 = view(Q);
        std::vector<uint32_t> tr = R.orderbuf();
        std::vector<double> rx(tr.size()), ry(tr.size()), rz(tr.size());
        gz_shmem::accel_grouped(R, Q, tr, 0.6, G, 0, ax, ay, az);
        gzo_accel_grouped(tq, &hq2, tr.data(), tr.size(), 0.6, G, 0, nullptr, rx.data(),
                          ry.data(), rz.data(), nullptr);
        e = max_rel(ax, ay, az, rx, ry, rz);
        expect(e < 1e-10, fmt("walk over refreshed nodes on drifted particles equals the oracle",
                              e));
        // raise_soft on every 50th particle
        std::vector<uint32_t> grow;
        for (size_t i = 0; i < m2; i += 50) grow.push_back((uint32_t)i);
        for (uint32_t i : grow) Q.soft[i] *= 3.0;
        R.raise_soft(grow, Q.soft);
        gzo_raise_soft(tq, grow.data(), grow.size(), Q.soft.data());
        const gzo_particles hq3 = view(Q);
        gz_shmem::accel_grouped(R, Q, tr, 0.6, G, 0, ax, ay, az);
        gzo_accel_grouped(tq, &hq3, tr.data(), tr.size(), 0.6, G, 0, nullptr, rx.data(),
                          ry.data(), rz.data(), nullptr);
        e = max_rel(ax, ay, az, rx, ry, rz);
        expect(e < 1e-10, fmt("walk after raise_soft equals the oracle", e));
        // neighbour search around a few particles, radius 0.2, on the aged tree
        std::vector<gz_shmem::Vec3d> cen;
        std::vector<double> rad, hcx, hcy, hcz;
        for (size_t i = 0; i < m2; i += 997) {
            cen.push_back(gz_shmem::Vec3d{{Q.x[i], Q.y[i], Q.z[i]}});
            rad.push_back(0.2);
            hcx.push_back(Q.x[i]);
            hcy.push_back(Q.y[i]);
            hcz.push_back(Q.z[i]);
        }
        std::vector<std::vector<uint32_t>> found;
        gz_shmem::ngb_search_many(R, Q, cen, rad, found);
        std::vector<size_t> ostart(cen.size() + 1);
        const size_t tot = gzo_ngb_search(tq, &hq3, hcx.data(), hcy.data(), hcz.data(),
                                          rad.data(), cen.size(), 0.1, ostart.data(), nullptr, 0);
        std::vector<uint32_t> ofound(tot + 1);
        gzo_ngb_search(tq, &hq3, hcx.data(), hcy.data(), hcz.data(), rad.data(), cen.size(), 0.1,
                       ostart.data(), ofound.data(), tot);
        bool same = found.size() == cen.size();
        size_t nfound = 0;
        for (size_t q = 0; same && q < cen.size(); ++q) {
            std::vector<uint32_t> a(found[q]), b(ofound.begin() + (std::ptrdiff_t)ostart[q],
                                                 ofound.begin() + (std::ptrdiff_t)ostart[q + 1]);
            std::sort(a.begin(), a.end());
            std::sort(b.begin(), b.end());
            same = a == b;
            nfound += a.size();
        }
        expect(same && nfound > 0, "neighbour lists equal the oracle's (" +
                                       std::to_string(nfound) + " neighbours)");
        std::vector<uint32_t> one;
        gz_shmem::ngb_search(R, Q, cen[0], rad[0], one);
        std::sort(one.begin(), one.end());
        std::vector<uint32_t> first(found[0]);
        std::sort(first.begin(), first.end());
        expect(one == first, "single-centre ngb_search appends the same list");
        // reverse search: kernel radii 0.05..0.25, the first 70% of the particles gas
        std::vector<double> hk(m2);
        for (size_t i = 0; i < m2; ++i) hk[i] = 0.05 + 0.2 * (double)((i * 7919) % 1000) / 1000.0;
        const size_t n_gas = (7 * m2) / 10;
        std::vector<double> hqr(cen.size(), 0.06);
        R.init_hmax(hk, n_gas);
        gzo_init_hmax(tq, hk.data(), hk.size(), n_gas, nullptr);
        std::vector<std::vector<uint32_t>> rfound;
        gz_shmem::ngb_search_reverse_many(R, Q, hk.data(), n_gas, cen, hqr, rfound);
        const size_t rtot = gzo_ngb_search_reverse(tq, &hq3, hk.data(), n_gas, hcx.data(),
                                                   hcy.data(), hcz.data(), hqr.data(), cen.size(),
                                                   0.1, ostart.data(), nullptr, 0);
        std::vector<uint32_t> rofound(rtot + 1);
        gzo_ngb_search_reverse(tq, &hq3, hk.data(), n_gas, hcx.data(), hcy.data(), hcz.data(),
                               hqr.data(), cen.size(), 0.1, ostart.data(), rofound.data(), rtot);
        bool rsame = rfound.size() == cen.size();
        size_t rn = 0;
        for (size_t q = 0; rsame && q < cen.size(); ++q) {
            const std::vector<uint32_t> b(rofound.begin() + (std::ptrdiff_t)ostart[q],
                                          rofound.begin() + (std::ptrdiff_t)ostart[q + 1]);
            rsame = rfound[q] == b;
            rn += b.size();
        }
        expect(rsame && rn > 0, "reverse neighbour lists equal the oracle's, in order (" +
                                    std::to_string(rn) + " pairs)");
        // kernels that grow between rebuilds: raise_hmax, then the wider reverse reach
        std::vector<uint32_t> hgrow;
        for (size_t i = 0; i < n_gas; i += 7) hgrow.push_back((uint32_t)i);
        for (uint32_t i : hgrow) hk[i] *= 2.0;
        R.raise_hmax(hgrow, hk);
        gzo_raise_hmax(tq, hgrow.data(), hgrow.size(), hk.data());
        std::vector<std::vector<uint32_t>> gfound;
        gz_shmem::ngb_search_reverse_many(R, Q, hk.data(), n_gas, cen, hqr, gfound);
        const size_t gtot = gzo_ngb_search_reverse(tq, &hq3, hk.data(), n_gas, hcx.data(),
                                                   hcy.data(), hcz.data(), hqr.data(), cen.size(),
                                                   0.1, ostart.data(), nullptr, 0);
        std::vector<uint32_t> gofound(gtot + 1);
        gzo_ngb_search_reverse(tq, &hq3, hk.data(), n_gas, hcx.data(), hcy.data(), hcz.data(),
                               hqr.data(), cen.size(), 0.1, ostart.data(), gofound.data(), gtot);
        bool gsame = gfound.size() == cen.size();
        for (size_t q = 0; gsame && q < cen.size(); ++q)
            gsame = gfound[q] == std::vector<uint32_t>(
                                     gofound.begin() + (std::ptrdiff_t)ostart[q],
                                     gofound.begin() + (std::ptrdiff_t)ostart[q + 1]);
        expect(gsame && gtot > rn, "after raise_hmax the reverse lists equal the oracle's (" +
                                       std::to_string(gtot) + " pairs)");
        gzo_tree_free(tq);
    }

    // ---- the density solve (hydro.cc:150-327) against the oracle's (pinned to the reference),
    // with some neighbours not gas, and its neighbour cache against the oracle's search at the
    // final h
    {
        gz_shmem::Particles D = plummer(6000, 1.0, 0.01, 777u);
        D.type.assign(D.size(), 0);
        for (size_t i = 0; i < D.size(); i += 9) D.type[i] = 4;
        const gzo_particles hd = view(D);
        void* td = gzo_build(&hd, -1);
        gz_shmem::Tree TD = gz_shmem::build(D);
        std::vector<uint32_t> dtg;
        for (uint32_t i = 0; i < (uint32_t)D.size(); i += 3) dtg.push_back(i);
        const size_t nd = dtg.size();
        gz_shmem::NeighborCache cache;
        const gz_shmem::DensityResult r =
            gz_shmem::density(TD, D, dtg, 32.0, 0.5, std::vector<double>(), 0.0, 3, &cache);
        std::vector<double> oh(nd), orho(nd);
        std::vector<int> on(nd), oit(nd);
        gzo_density(td, &hd, D.type.data(), dtg.data(), nd, 32.0, 0.5, nullptr, 3, nullptr, 0.0,
                    oh.data(), orho.data(), on.data(), oit.data());
        double eh = 0.0;
        bool counts = true;
        for (size_t k = 0; k < nd; ++k) {
            eh = std::fmax(eh, std::fabs(r.h[k] - oh[k]) / oh[k]);
            eh = std::fmax(eh, std::fabs(r.rho[k] - orho[k]) / orho[k]);
            counts = counts && r.nngb[k] == on[k] && r.iters[k] == oit[k];
        }
        expect(counts && eh <= 1e-12,
               fmt("density: iterations and counts equal the oracle's, h and rho within", eh));
        std::vector<double> cx(nd), cy(nd), cz(nd);
        for (size_t k = 0; k < nd; ++k) {
            cx[k] = D.x[dtg[k]];
            cy[k] = D.y[dtg[k]];
            cz[k] = D.z[dtg[k]];
        }
        std::vector<size_t> ostart(nd + 1);
        const size_t tot = gzo_ngb_search(td, &hd, cx.data(), cy.data(), cz.data(), oh.data(), nd,
                                          0.0, ostart.data(), nullptr, 0);
        std::vector<uint32_t> oflat(tot);
        gzo_ngb_search(td, &hd, cx.data(), cy.data(), cz.data(), oh.data(), nd, 0.0,
                       ostart.data(), oflat.data(), tot);
        expect(cache.valid && cache.start == ostart && cache.flat == oflat,
               "density: the neighbour cache equals the oracle's search at the final h (" +
                   std::to_string(tot) + " pairs)");
        gzo_tree_free(td);
    }

    std::printf("%s: %d failure(s)\n", g_failures ? "shim FAILED" : "shim ok", g_failures);
    return g_failures ? 1 : 0;
}
