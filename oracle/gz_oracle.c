/* CPU oracle for the tree-gravity path: a plain-C restatement of the reference shared-memory tree.
 * TEST INFRASTRUCTURE ONLY (tests, smoke check, CPU-baseline leg of the benchmark); the product
 * library never links it. How each function is pinned is listed in gz_oracle.h.
 *
 * Compiled with -ffp-contract=off (oracle/Makefile): every expression below is evaluated in the
 * order written, with no fused multiply-add, so results are bit-comparable with the reference
 * built the same way. Where the reference's floating-point expression order matters for bit
 * equality (spline kernel, opening test, moment sums) it is kept as the reference writes it. */
#include "gz_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LEAF_MAX 16  /* tree.h:357 */
#define MAX_LEVEL 20 /* tree.h:358 */

static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* Spread the low 21 bits of v so bit k lands at bit 3k (tree.h:162-170), here as a loop over the
 * bits rather than the reference's mask cascade; the two are compared in the tests. */
static uint64_t spread_bits3(uint64_t v) {
    uint64_t out = 0;
    for (int k = 0; k < 21; ++k) out |= ((v >> k) & 1ull) << (3 * k);
    return out;
}

/* Morton key, x in bit 0, y in bit 1, z in bit 2 (tree.h:171-173). */
uint64_t gzo_morton(uint32_t a, uint32_t b, uint32_t c) {
    return spread_bits3(a) | (spread_bits3(b) << 1) | (spread_bits3(c) << 2);
}

/* Cubic-spline force/r, h the kernel extent (tree.h:59-67). */
double gzo_spline_force_over_r(double r, double h) {
    const double hi = 1.0 / h;
    const double hi3 = hi * hi * hi;
    const double u = r * hi;
    if (u >= 1.0) return 1.0 / (r * r * r);
    if (u < 0.5) return hi3 * (10.666666666666667 + u * u * (32.0 * u - 38.4));
    return hi3 * (21.333333333333332 - 48.0 * u + 38.4 * u * u - 10.666666666666667 * u * u * u -
                  0.06666666666666667 / (u * u * u));
}

/* Cubic-spline potential (tree.h:72-80). */
double gzo_spline_potential(double r, double h) {
    const double hi = 1.0 / h;
    const double u = r * hi;
    if (u >= 1.0) return -1.0 / r;
    if (u < 0.5) return hi * (-2.8 + u * u * (5.333333333333333 + u * u * (6.4 * u - 9.6)));
    return hi * (-3.2 + 0.06666666666666667 / u +
                 u * u * (10.666666666666667 + u * (-16.0 + u * (9.6 - 2.1333333333333333 * u))));
}

/* Node-opening decision (tree.h:109-144): softening overlap, Barnes-Hut when theta_sq > 0,
 * relative criterion when aold > 0 (open everything when neither criterion is live), and the
 * inside-node test. Returns 1 to open. */
int gzo_open_node(double r_sq, double len, double l_open, double node_mass, double node_maxsoft,
                  double soft_target, double aold, double theta_sq, double adx, double ady,
                  double adz) {
    const double reach_t = soft_target + 0.6 * len;
    const double reach_n = node_maxsoft + 0.6 * len;
    if (r_sq < reach_t * reach_t || r_sq < reach_n * reach_n) return 1;
    if (theta_sq > 0.0 && l_open * l_open >= theta_sq * r_sq) return 1;
    if (aold > 0) {
        if (node_mass * len * len > r_sq * r_sq * aold) return 1;
    } else if (theta_sq <= 0.0) {
        return 1;
    }
    return (adx < 0.6 * len && ady < 0.6 * len && adz < 0.6 * len) ? 1 : 0;
}

/* ---------------------------------------------------------------------------------------------
 * Keys (tree.cc:355-418)
 * ------------------------------------------------------------------------------------------- */

/* splitmix64 stream for RANDOMIZE_GRAVTREE (tree.cc:381-393). */
typedef struct {
    uint64_t s;
} splitmix;
static double splitmix_unit(splitmix* g) {
    uint64_t z = (g->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

/* shift: SHMEM_TREE_SHIFT's three fractions (tree.cc:400-407), or NULL */
static void root_box(const gzo_particles* P, long long seed, const double* shift, double* root) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < P->n; ++i) {
        if (!(P->m[i] > 0)) continue;
        lo[0] = dmin(lo[0], P->x[i]); hi[0] = dmax(hi[0], P->x[i]);
        lo[1] = dmin(lo[1], P->y[i]); hi[1] = dmax(hi[1], P->y[i]);
        lo[2] = dmin(lo[2], P->z[i]); hi[2] = dmax(hi[2], P->z[i]);
    }
    double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    double side = dmax(hi[0] - lo[0], dmax(hi[1] - lo[1], hi[2] - lo[2])) * 1.0000001;
    if (side <= 0) side = 1.0;
    if (seed >= 0) {
        splitmix g = {(uint64_t)seed * 0x9E3779B97F4A7C15ull};
        cx += side * (splitmix_unit(&g) - 0.5);
        cy += side * (splitmix_unit(&g) - 0.5);
        cz += side * (splitmix_unit(&g) - 0.5);
        side *= 2.0;
    }
    if (shift) {
        cx += shift[0] * side; cy += shift[1] * side; cz += shift[2] * side;
        side *= 1.0 + 2.0 * (fabs(shift[0]) + fabs(shift[1]) + fabs(shift[2]));
    }
    root[0] = cx; root[1] = cy; root[2] = cz; root[3] = side;
}

static inline uint64_t key_of(double x, double y, double z, const double* root) {
    const double scale = (double)((1u << MAX_LEVEL) - 1) / root[3];
    const double half = 0.5 * root[3];
    const uint32_t a = (uint32_t)((x - (root[0] - half)) * scale);
    const uint32_t b = (uint32_t)((y - (root[1] - half)) * scale);
    const uint32_t c = (uint32_t)((z - (root[2] - half)) * scale);
    return gzo_morton(a, b, c);
}

static const double* g_shift = NULL; /* the shift of the build in progress (gzo_build_shift) */

void gzo_keys(const gzo_particles* P, long long seed, uint64_t* keys, double* root) {
    root_box(P, seed, g_shift, root);
    for (size_t i = 0; i < P->n; ++i)
        keys[i] = (P->m[i] > 0) ? key_of(P->x[i], P->y[i], P->z[i], root) : 0ull;
}

/* ---------------------------------------------------------------------------------------------
 * Build (tree.cc:70-197, 340-348, 350-526)
 * ------------------------------------------------------------------------------------------- */

typedef struct {
    int nnodes, cap, root;
    size_t n, n_live;
    int *first, *next, *plo, *phi, *parent;
    double *size, *mass, *cx, *cy, *cz, *soft, *delta, *gcx, *gcy, *gcz, *s;
    float *wsoft, *wlen;
    /* the walk's view of each node (WNode, tree.h:195-204): COM and mass, rewritten by
     * gzo_refresh_walk_nodes; the build moments above stay as built */
    double *wcx, *wcy, *wcz, *wmass;
    float* vmax; /* max |v| below the node since the build (0 without velocities) */
    float* xpad;
    float* hmax; /* largest gas kernel radius below the node (gzo_init_hmax), or NULL */
    /* with velocities: centre-of-mass velocity and the incremental moments (tree.h:300-317) */
    double *vcom_x, *vcom_y, *vcom_z, *dp_x, *dp_y, *dp_z, *dq_x, *dq_y, *dq_z, *dmass;
    uint32_t *orderbuf, *rank;
    int* leaf_of;
    uint64_t* skey;
    const double *vx, *vy, *vz; /* build-time velocities (only during the build) */
} otree;

static void grow(otree* T) {
    int cap = T->cap ? 2 * T->cap : 1024;
#define RE(f, ty) T->f = (ty*)realloc(T->f, (size_t)cap * sizeof(ty))
    RE(first, int); RE(next, int); RE(plo, int); RE(phi, int); RE(parent, int);
    RE(size, double); RE(mass, double); RE(cx, double); RE(cy, double); RE(cz, double);
    RE(soft, double); RE(delta, double); RE(gcx, double); RE(gcy, double); RE(gcz, double);
    RE(s, double); RE(wsoft, float); RE(wlen, float); RE(vmax, float);
    if (T->vx) {
        RE(vcom_x, double); RE(vcom_y, double); RE(vcom_z, double);
    }
#undef RE
    T->cap = cap;
}

typedef struct {
    uint64_t key;
    uint32_t idx;
} keyed;

static int keyed_cmp(const void* a, const void* b) {
    const keyed* p = (const keyed*)a;
    const keyed* q = (const keyed*)b;
    if (p->key != q->key) return (p->key < q->key) ? -1 : 1;
    return (p->idx < q->idx) ? -1 : (p->idx > q->idx);
}

/* One node over the sorted range [lo, hi) at `level`, geometric centre (gx,gy,gz), side `side`.
 * Node ids are claimed in preorder. Children are the non-empty octants in octant order; their
 * centres come from halving this node's box (tree.cc:101, 116-118), not from key bits. */
static int make_node(otree* T, const gzo_particles* P, int lo, int hi, int level, double gx,
                     double gy, double gz, double side) {
    if (T->nnodes == T->cap) grow(T);
    const int me = T->nnodes++;
    T->size[me] = side; T->first[me] = -1; T->next[me] = -1;
    T->plo[me] = lo; T->phi[me] = hi; T->parent[me] = -1;
    T->gcx[me] = gx; T->gcy[me] = gy; T->gcz[me] = gz;

    int kids[8], nk = 0;
    if (!((hi - lo <= LEAF_MAX) || (level >= MAX_LEVEL))) {
        const int shift = 3 * (MAX_LEVEL - level - 1);
        const double half = side * 0.5, quarter = side * 0.25;
        int at = lo;
        for (int oct = 0; oct < 8; ++oct) {
            const int start = at;
            while (at < hi && (int)((T->skey[at] >> shift) & 7ull) == oct) ++at;
            if (at == start) continue;
            const double ox = gx + ((oct & 1) ? quarter : -quarter);
            const double oy = gy + ((oct & 2) ? quarter : -quarter);
            const double oz = gz + ((oct & 4) ? quarter : -quarter);
            kids[nk++] = make_node(T, P, start, at, level + 1, ox, oy, oz, half);
        }
        for (int i = 0; i < nk; ++i) T->parent[kids[i]] = me;
        T->first[me] = kids[0];
        for (int i = 0; i + 1 < nk; ++i) T->next[kids[i]] = kids[i + 1];
    }

    /* Moments: a leaf sums its particles in sorted order; an interior node combines its
     * children's final moments in octant order (tree.cc:153-190). */
    double M = 0, sx = 0, sy = 0, sz = 0, smax = 0;
    double vmax = 0, pvx = 0, pvy = 0, pvz = 0; /* with velocities (tree.cc:161-166, 177-181) */
    if (nk == 0) {
        for (int i = lo; i < hi; ++i) {
            const uint32_t p = T->orderbuf[i];
            const double m = P->m[p];
            M += m; sx += m * P->x[p]; sy += m * P->y[p]; sz += m * P->z[p];
            if (P->soft && P->soft[p] > smax) smax = P->soft[p];
            if (T->vx) {
                const double v = sqrt(T->vx[p] * T->vx[p] + T->vy[p] * T->vy[p] +
                                      T->vz[p] * T->vz[p]);
                if (v > vmax) vmax = v;
                pvx += m * T->vx[p]; pvy += m * T->vy[p]; pvz += m * T->vz[p];
            }
            T->leaf_of[p] = me;
        }
    } else {
        for (int i = 0; i < nk; ++i) {
            const int k = kids[i];
            const double m = T->mass[k];
            M += m; sx += m * T->cx[k]; sy += m * T->cy[k]; sz += m * T->cz[k];
            if (T->soft[k] > smax) smax = T->soft[k];
            if (T->vmax[k] > vmax) vmax = T->vmax[k];
            if (T->vx) { pvx += m * T->vcom_x[k]; pvy += m * T->vcom_y[k]; pvz += m * T->vcom_z[k]; }
        }
    }
    if (M > 0) { sx /= M; sy /= M; sz /= M; }
    T->mass[me] = M; T->cx[me] = sx; T->cy[me] = sy; T->cz[me] = sz; T->soft[me] = smax;
    T->vmax[me] = (float)vmax;
    if (T->vx) {
        T->vcom_x[me] = 0; T->vcom_y[me] = 0; T->vcom_z[me] = 0;
        if (M > 0) { T->vcom_x[me] = pvx / M; T->vcom_y[me] = pvy / M; T->vcom_z[me] = pvz / M; }
    }
    const double dx = sx - gx, dy = sy - gy, dz = sz - gz;
    T->delta[me] = sqrt(dx * dx + dy * dy + dz * dz);
    return me;
}

/* next[] threading: the last child of a node continues at its parent's successor (tree.cc:340). */
static void thread_links(otree* T, int node, int successor) {
    T->next[node] = successor;
    int c = T->first[node];
    while (c >= 0) {
        const int sib = T->next[c];
        thread_links(T, c, sib >= 0 ? sib : successor);
        c = sib;
    }
}

void* gzo_build(const gzo_particles* P, long long seed) {
    return gzo_build_vel(P, seed, NULL, NULL, NULL);
}

void* gzo_build_shift(const gzo_particles* P, long long seed, double fx, double fy, double fz) {
    const double sh[3] = {fx, fy, fz};
    g_shift = sh;
    void* t = gzo_build_vel(P, seed, NULL, NULL, NULL);
    g_shift = NULL;
    return t;
}

void* gzo_build_vel(const gzo_particles* P, long long seed, const double* vx, const double* vy,
                    const double* vz) {
    otree* T = (otree*)calloc(1, sizeof(otree));
    if (vx && vy && vz) { T->vx = vx; T->vy = vy; T->vz = vz; }
    const size_t n = P->n;
    T->n = n;
    double root[4];
    uint64_t* keys = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    gzo_keys(P, seed, keys, root);

    /* live particles in index order, then sorted by key; ties keep index order */
    size_t nl = 0;
    keyed* kv = (keyed*)malloc((n ? n : 1) * sizeof(keyed));
    for (size_t i = 0; i < n; ++i)
        if (P->m[i] > 0) { kv[nl].key = keys[i]; kv[nl].idx = (uint32_t)i; ++nl; }
    qsort(kv, nl, sizeof(keyed), keyed_cmp);
    T->n_live = nl;
    T->orderbuf = (uint32_t*)malloc((nl ? nl : 1) * sizeof(uint32_t));
    T->skey = (uint64_t*)malloc((nl ? nl : 1) * sizeof(uint64_t));
    for (size_t r = 0; r < nl; ++r) { T->orderbuf[r] = kv[r].idx; T->skey[r] = kv[r].key; }
    free(kv);
    free(keys);

    T->leaf_of = (int*)malloc((n ? n : 1) * sizeof(int));
    for (size_t i = 0; i < n; ++i) T->leaf_of[i] = -1;
    T->root = make_node(T, P, 0, (int)nl, 0, root[0], root[1], root[2], root[3]);
    thread_links(T, T->root, -1);

    T->rank = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) T->rank[i] = UINT32_MAX;
    for (size_t r = 0; r < nl; ++r) T->rank[T->orderbuf[r]] = (uint32_t)r;

    /* packed walk fields (tree.cc:507-523) */
    const size_t nn = (size_t)(T->nnodes ? T->nnodes : 1);
    T->wcx = (double*)malloc(nn * sizeof(double)); T->wcy = (double*)malloc(nn * sizeof(double));
    T->wcz = (double*)malloc(nn * sizeof(double));
    T->wmass = (double*)malloc(nn * sizeof(double));
    T->xpad = (float*)calloc(nn, sizeof(float));
    for (int i = 0; i < T->nnodes; ++i) {
        T->s[i] = T->size[i] + T->delta[i];
        T->wsoft[i] = (float)T->soft[i];
        T->wlen[i] = (float)T->size[i];
        T->wcx[i] = T->cx[i]; T->wcy[i] = T->cy[i]; T->wcz[i] = T->cz[i];
        T->wmass[i] = T->mass[i];
    }
    if (T->vx) { /* incremental moments start at zero (tree.cc:449-453) */
        T->dp_x = (double*)calloc(nn, sizeof(double)); T->dp_y = (double*)calloc(nn, sizeof(double));
        T->dp_z = (double*)calloc(nn, sizeof(double)); T->dq_x = (double*)calloc(nn, sizeof(double));
        T->dq_y = (double*)calloc(nn, sizeof(double)); T->dq_z = (double*)calloc(nn, sizeof(double));
        T->dmass = (double*)calloc(nn, sizeof(double));
    }
    T->vx = T->vy = T->vz = NULL;
    return T;
}

void gzo_tree_free(void* h) {
    otree* T = (otree*)h;
    if (!T) return;
    free(T->first); free(T->next); free(T->plo); free(T->phi); free(T->parent);
    free(T->size); free(T->mass); free(T->cx); free(T->cy); free(T->cz); free(T->soft);
    free(T->delta); free(T->gcx); free(T->gcy); free(T->gcz); free(T->s); free(T->wsoft);
    free(T->wlen); free(T->orderbuf); free(T->rank); free(T->leaf_of); free(T->skey);
    free(T->wcx); free(T->wcy); free(T->wcz); free(T->wmass); free(T->vmax); free(T->xpad);
    free(T->hmax);
    free(T->vcom_x); free(T->vcom_y); free(T->vcom_z); free(T->dp_x); free(T->dp_y);
    free(T->dp_z); free(T->dq_x); free(T->dq_y); free(T->dq_z); free(T->dmass);
    free(T);
}

void gzo_tree_sizes(void* h, long long* out) {
    otree* T = (otree*)h;
    out[0] = T->nnodes; out[1] = T->root; out[2] = (long long)T->n_live;
}

void gzo_tree_nodes(void* h, int* first, int* next, int* plo, int* phi, int* parent,
                    double* size, double* mass, double* cx, double* cy, double* cz, double* soft,
                    double* delta, double* gcx, double* gcy, double* gcz, double* wn_s,
                    float* wn_soft, float* wn_len) {
    otree* T = (otree*)h;
    const size_t nn = (size_t)T->nnodes;
#define CP(dst, src, ty) if (dst) memcpy(dst, T->src, nn * sizeof(ty))
    CP(first, first, int); CP(next, next, int); CP(plo, plo, int); CP(phi, phi, int);
    CP(parent, parent, int); CP(size, size, double); CP(mass, mass, double);
    CP(cx, cx, double); CP(cy, cy, double); CP(cz, cz, double); CP(soft, soft, double);
    CP(delta, delta, double); CP(gcx, gcx, double); CP(gcy, gcy, double); CP(gcz, gcz, double);
    CP(wn_s, s, double); CP(wn_soft, wsoft, float); CP(wn_len, wlen, float);
#undef CP
}

void gzo_tree_particles(void* h, uint32_t* orderbuf, uint32_t* rank, int* leaf_of) {
    otree* T = (otree*)h;
    if (orderbuf) memcpy(orderbuf, T->orderbuf, T->n_live * sizeof(uint32_t));
    if (rank) memcpy(rank, T->rank, T->n * sizeof(uint32_t));
    if (leaf_of) memcpy(leaf_of, T->leaf_of, T->n * sizeof(int));
}

void gzo_tree_skeys(void* h, uint64_t* skey) {
    otree* T = (otree*)h;
    memcpy(skey, T->skey, T->n_live * sizeof(uint64_t));
}

/* ---------------------------------------------------------------------------------------------
 * Tree reuse between rebuilds (tree.cc:211-274, 276-285, 311-319; the kick feed mfm.cc:3798-3811)
 * ------------------------------------------------------------------------------------------- */

/* The feed of one step's velocity changes, in the order of `active`: raise every ancestor's vmax
 * to |v| (stopping at the first that already covers it, tree.cc:276-285), add the momentum change
 * m (v - v_last) and its first moment times tau up the parents (tree.cc:211-227), then
 * v_last = v. Without velocities at build time only vmax is raised; with vlast NULL no momentum is
 * fed. Sequential: the result is that of the reference run on one thread. */
void gzo_kick_nodes(void* h, const uint32_t* active, size_t na, const double* vx, const double* vy,
                    const double* vz, const double* m, double* vlx, double* vly, double* vlz,
                    double tau) {
    otree* T = (otree*)h;
    for (size_t k = 0; k < na; ++k) {
        const uint32_t i = active[k];
        const int leaf = T->leaf_of[i];
        if (leaf < 0) continue;
        const double vn = sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        const float speed = (float)vn;
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            if (T->vmax[no] >= speed) break;
            T->vmax[no] = speed;
        }
        if (T->dp_x && vlx) {
            const double px = (vx[i] - vlx[i]) * m[i], py = (vy[i] - vly[i]) * m[i],
                         pz = (vz[i] - vlz[i]) * m[i];
            for (int no = leaf; no >= 0; no = T->parent[no]) {
                T->dp_x[no] += px; T->dp_y[no] += py; T->dp_z[no] += pz;
                T->dq_x[no] += px * tau; T->dq_y[no] += py * tau; T->dq_z[no] += pz * tau;
            }
        }
        if (vlx) { vlx[i] = vx[i]; vly[i] = vy[i]; vlz[i] = vz[i]; }
    }
}

/* refresh_walk_nodes (tree.cc:260-274) with node_mass / node_com (tree.h:318-320, tree.cc:230-237):
 * M = mass + dmass, COM = (m0 C + (m0 vcom + dp) tau - dq) / M, opening radius and side widened
 * by 2 (vmax tau + xpad). */
void gzo_refresh_walk_nodes(void* h, double tau) {
    otree* T = (otree*)h;
    for (int i = 0; i < T->nnodes; ++i) {
        const double M = T->mass[i] + (T->dmass ? T->dmass[i] : 0.0);
        double c0 = T->cx[i], c1 = T->cy[i], c2 = T->cz[i];
        if (T->dp_x && M > 0) {
            const double m0 = T->mass[i];
            c0 = (m0 * T->cx[i] + (m0 * T->vcom_x[i] + T->dp_x[i]) * tau - T->dq_x[i]) / M;
            c1 = (m0 * T->cy[i] + (m0 * T->vcom_y[i] + T->dp_y[i]) * tau - T->dq_y[i]) / M;
            c2 = (m0 * T->cz[i] + (m0 * T->vcom_z[i] + T->dp_z[i]) * tau - T->dq_z[i]) / M;
        }
        T->wcx[i] = c0; T->wcy[i] = c1; T->wcz[i] = c2;
        T->wmass[i] = M;
        const double grow = 2.0 * ((double)T->vmax[i] * tau + (double)T->xpad[i]);
        T->s[i] = T->size[i] + T->delta[i] + grow;
        T->wlen[i] = (float)(T->size[i] + grow);
    }
}

/* soft rounded up to float, so the node bound stays one-sided (tree.cc:289-292). */
static float soft_as_bound(double h) {
    float f = (float)h;
    return ((double)f < h) ? nextafterf(f, 1e30f) : f;
}

/* raise_soft (tree.cc:311-319) for each listed particle, in order. */
void gzo_raise_soft(void* h, const uint32_t* idx, size_t n, const double* soft) {
    otree* T = (otree*)h;
    for (size_t k = 0; k < n; ++k) {
        const uint32_t i = idx[k];
        const int leaf = T->leaf_of[i];
        if (leaf < 0) continue;
        const float sf = soft_as_bound(soft[i]);
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            if (T->wsoft[no] >= sf) break;
            T->wsoft[no] = sf;
            if (T->soft[no] < soft[i]) T->soft[no] = soft[i];
        }
    }
}

/* The minimum-image separation of vec.h:25-30: folded into [-box/2, box/2] when box > 0. */
static double min_image_sep(double d, double box) {
    if (box > 0) {
        if (d > 0.5 * box) d -= box;
        else if (d < -0.5 * box) d += box;
    }
    return d;
}

/* node_com (tree.cc:229-237) at tau from the incremental moments: the build COM when the node
 * has no mass left. Only called on trees built with velocities. */
static void node_com_at(const otree* T, int no, double tau, double* c) {
    const double M = T->mass[no] + T->dmass[no];
    if (M <= 0) { c[0] = T->cx[no]; c[1] = T->cy[no]; c[2] = T->cz[no]; return; }
    const double m0 = T->mass[no];
    c[0] = (m0 * T->cx[no] + (m0 * T->vcom_x[no] + T->dp_x[no]) * tau - T->dq_x[no]) / M;
    c[1] = (m0 * T->cy[no] + (m0 * T->vcom_y[no] + T->dp_y[no]) * tau - T->dq_y[no]) / M;
    c[2] = (m0 * T->cz[no] + (m0 * T->vcom_z[no] + T->dp_z[no]) * tau - T->dq_z[no]) / M;
}

/* mass_event (tree.cc:239-251) for each event in order: every ancestor of the leaf takes dm, its
 * momentum dm v and, in its first moment, dm (x_r - v tau), where x_r is x unwrapped against that
 * ancestor's COM at the moment of the event. */
void gzo_mass_events(void* h, const uint32_t* idx, size_t n, const double* dm, const double* x,
                     const double* y, const double* z, const double* vx, const double* vy,
                     const double* vz, double tau, double box) {
    otree* T = (otree*)h;
    if (!T->dp_x) return;
    for (size_t k = 0; k < n; ++k) {
        const int leaf = T->leaf_of[idx[k]];
        if (leaf < 0 || dm[k] == 0.0) continue;
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            double c[3];
            node_com_at(T, no, tau, c);
            const double xr0 = c[0] + min_image_sep(x[k] - c[0], box);
            const double xr1 = c[1] + min_image_sep(y[k] - c[1], box);
            const double xr2 = c[2] + min_image_sep(z[k] - c[2], box);
            T->dmass[no] += dm[k];
            T->dp_x[no] += dm[k] * vx[k];
            T->dp_y[no] += dm[k] * vy[k];
            T->dp_z[no] += dm[k] * vz[k];
            T->dq_x[no] -= dm[k] * (xr0 - vx[k] * tau);
            T->dq_y[no] -= dm[k] * (xr1 - vy[k] * tau);
            T->dq_z[no] -= dm[k] * (xr2 - vz[k] * tau);
        }
    }
}

/* shift_node (tree.cc:253-258): every ancestor's first moment loses m dx. */
void gzo_shift_nodes(void* h, const uint32_t* idx, size_t n, const double* m, const double* dx,
                     const double* dy, const double* dz) {
    otree* T = (otree*)h;
    if (!T->dp_x) return;
    for (size_t k = 0; k < n; ++k) {
        const int leaf = T->leaf_of[idx[k]];
        if (leaf < 0) continue;
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            T->dq_x[no] -= m[k] * dx[k];
            T->dq_y[no] -= m[k] * dy[k];
            T->dq_z[no] -= m[k] * dz[k];
        }
    }
}

/* raise_pad (tree.cc:323-330): the leaf's pad grows by d (rounded up to float) and every ancestor
 * takes it as a maximum, stopping at the first that already covers it. */
void gzo_raise_pad(void* h, const uint32_t* idx, size_t n, const double* d) {
    otree* T = (otree*)h;
    for (size_t k = 0; k < n; ++k) {
        const int leaf = T->leaf_of[idx[k]];
        if (leaf < 0 || !(d[k] > 0)) continue;
        const float pad = soft_as_bound((double)T->xpad[leaf] + d[k]);
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            if (T->xpad[no] >= pad) break;
            T->xpad[no] = pad;
        }
    }
}

/* raise_vmax (tree.cc:276-285) to speed[k], stopping at the first ancestor that covers it. */
void gzo_raise_vmax(void* h, const uint32_t* idx, size_t n, const float* speed) {
    otree* T = (otree*)h;
    for (size_t k = 0; k < n; ++k) {
        const int leaf = T->leaf_of[idx[k]];
        if (leaf < 0) continue;
        for (int no = leaf; no >= 0; no = T->parent[no]) {
            if (T->vmax[no] >= speed[k]) break;
            T->vmax[no] = speed[k];
        }
    }
}

void gzo_tree_pads(void* h, float* xpad, double* dmass) {
    const otree* T = (const otree*)h;
    for (int i = 0; i < T->nnodes; ++i) {
        if (xpad) xpad[i] = T->xpad[i];
        if (dmass) dmass[i] = T->dmass ? T->dmass[i] : 0.0;
    }
}

/* The reuse state per node; any pointer may be NULL (vcom, dp and dq are zero without
 * velocities at build time). */
void gzo_tree_reuse(void* h, float* vmax, double* vcom, double* dp, double* dq, double* wcom,
                    double* wmass) {
    otree* T = (otree*)h;
    const int nn = T->nnodes;
    for (int i = 0; i < nn; ++i) {
        if (vmax) vmax[i] = T->vmax[i];
        if (vcom) {
            vcom[3 * i] = T->vcom_x ? T->vcom_x[i] : 0.0;
            vcom[3 * i + 1] = T->vcom_x ? T->vcom_y[i] : 0.0;
            vcom[3 * i + 2] = T->vcom_x ? T->vcom_z[i] : 0.0;
        }
        if (dp) {
            dp[3 * i] = T->dp_x ? T->dp_x[i] : 0.0;
            dp[3 * i + 1] = T->dp_x ? T->dp_y[i] : 0.0;
            dp[3 * i + 2] = T->dp_x ? T->dp_z[i] : 0.0;
        }
        if (dq) {
            dq[3 * i] = T->dq_x ? T->dq_x[i] : 0.0;
            dq[3 * i + 1] = T->dq_x ? T->dq_y[i] : 0.0;
            dq[3 * i + 2] = T->dq_x ? T->dq_z[i] : 0.0;
        }
        if (wcom) { wcom[3 * i] = T->wcx[i]; wcom[3 * i + 1] = T->wcy[i]; wcom[3 * i + 2] = T->wcz[i]; }
        if (wmass) wmass[i] = T->wmass[i];
    }
}

/* ---------------------------------------------------------------------------------------------
 * Neighbour search (hydro.cc:37-108) over the search node (SNode, tree.h:206-226, packed at
 * tree.cc:496-506): float geometric centre and a conservatively widened half side
 * ------------------------------------------------------------------------------------------- */

/* All particles within radius[q] of centre q (non-periodic), in traversal order, for a tree aged
 * tau. CSR: start[q] .. start[q+1]; found written while it fits in cap. Returns the total. */
size_t gzo_ngb_search(void* h, const gzo_particles* P, const double* cx, const double* cy,
                      const double* cz, const double* radius, size_t nq, double tau, size_t* start,
                      uint32_t* found, size_t cap) {
    const otree* T = (const otree*)h;
    const double snode_eps = 1e-6 * (T->nnodes ? T->size[T->root] : 1.0);
    const int nn = T->nnodes;
    float* sc = (float*)malloc((size_t)(nn ? nn : 1) * 4 * sizeof(float));
    for (int i = 0; i < nn; ++i) {
        sc[4 * i] = (float)T->gcx[i]; sc[4 * i + 1] = (float)T->gcy[i];
        sc[4 * i + 2] = (float)T->gcz[i];
        const double cmax = fmax(fabs(T->gcx[i]), fmax(fabs(T->gcy[i]), fabs(T->gcz[i])));
        sc[4 * i + 3] = (float)(0.5 * T->size[i] + snode_eps + 4.0 * (double)FLT_EPSILON * cmax);
    }
    size_t total = 0;
    start[0] = 0;
    for (size_t q = 0; q < nq; ++q) {
        const double r2max = radius[q] * radius[q];
        int node = nn ? T->root : -1;
        while (node >= 0) {
            const float* sn = sc + 4 * node;
            double dist = radius[q] + (double)sn[3] + (double)T->xpad[node];
            if (tau > 0) dist += (double)T->vmax[node] * tau;
            const double dx = (double)sn[0] - cx[q];
            if (dx > dist || -dx > dist) { node = T->next[node]; continue; }
            const double dy = (double)sn[1] - cy[q];
            if (dy > dist || -dy > dist) { node = T->next[node]; continue; }
            const double dz = (double)sn[2] - cz[q];
            if (dz > dist || -dz > dist) { node = T->next[node]; continue; }
            dist += 0.7320508075688772 * (double)sn[3];
            if (dx * dx + dy * dy + dz * dz > dist * dist) { node = T->next[node]; continue; }
            if (T->first[node] < 0) {
                for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                    const uint32_t j = T->orderbuf[k];
                    const double ex = P->x[j] - cx[q], ey = P->y[j] - cy[q], ez = P->z[j] - cz[q];
                    if (ex * ex + ey * ey + ez * ez < r2max) {
                        if (found && total < cap) found[total] = j;
                        ++total;
                    }
                }
                node = T->next[node];
            } else {
                node = T->first[node];
            }
        }
        start[q + 1] = total;
    }
    free(sc);
    return total;
}

/* h rounded up to float, so that a float bound stays one-sided (tree.cc:286-290). */
static float h_as_bound(double h) {
    float f = (float)h;
    return ((double)f < h) ? nextafterf(f, 1e30f) : f;
}

/* init_hmax (tree.cc:332-338, raise_hmax tree.cc:293-301): every node's hmax is the largest
 * rounded-up h of the gas particles (index < n_gas, h > 0) below it. The reference raises each
 * particle's leaf and ancestors with an early exit; the result is the same subtree maximum. */
void gzo_init_hmax(void* t, const double* h, size_t nh, size_t n_gas, float* hmax_out) {
    otree* T = (otree*)t;
    const int nn = T->nnodes;
    free(T->hmax);
    T->hmax = (float*)calloc((size_t)(nn ? nn : 1), sizeof(float));
    size_t n = n_gas < nh ? n_gas : nh;
    if (n > T->n) n = T->n;
    for (size_t i = 0; i < n; ++i) {
        if (!(h[i] > 0) || T->leaf_of[i] < 0) continue;
        const float hf = h_as_bound(h[i]);
        for (int no = T->leaf_of[i]; no >= 0; no = T->parent[no]) {
            if (T->hmax[no] >= hf) break;
            T->hmax[no] = hf;
        }
    }
    if (hmax_out) memcpy(hmax_out, T->hmax, (size_t)nn * sizeof(float));
}

/* The node hmax (nnodes floats; zeros before gzo_init_hmax). */
void gzo_tree_hmax(void* t, float* out) {
    const otree* T = (const otree*)t;
    for (int i = 0; i < T->nnodes; ++i) out[i] = T->hmax ? T->hmax[i] : 0.0f;
}

/* raise_hmax (tree.cc:293-301) for the leaf of each listed particle i, to h[i] (by particle);
 * h[i] <= 0 raises nothing. Needs gzo_init_hmax. */
void gzo_raise_hmax(void* t, const uint32_t* idx, size_t n, const double* h) {
    otree* T = (otree*)t;
    if (!T->hmax) return;
    for (size_t k = 0; k < n; ++k) {
        const uint32_t i = idx[k];
        if (i >= T->n || T->leaf_of[i] < 0 || !(h[i] > 0)) continue;
        const float hf = h_as_bound(h[i]);
        for (int no = T->leaf_of[i]; no >= 0; no = T->parent[no]) {
            if (T->hmax[no] >= hf) break;
            T->hmax[no] = hf;
        }
    }
}

/* ngb_search_reverse (hydro.cc:110-148), non-periodic, tree aged tau: for each centre q with
 * kernel radius hq[q], every gas particle j (index < n_gas) with hq[q] <= r < h[j], in traversal
 * (Morton) order. A node is skipped when no particle below it has h > hq[q] (hmax <= hq[q]) or
 * when it lies beyond hmax + half + xpad (+ vmax tau) of the centre. Needs gzo_init_hmax. */
size_t gzo_ngb_search_reverse(void* t, const gzo_particles* P, const double* h, size_t n_gas,
                              const double* cx, const double* cy, const double* cz,
                              const double* hq, size_t nq, double tau, size_t* start,
                              uint32_t* found, size_t cap) {
    const otree* T = (const otree*)t;
    const int nn = T->nnodes;
    start[0] = 0;
    if (!T->hmax) {
        for (size_t q = 0; q < nq; ++q) start[q + 1] = 0;
        return 0;
    }
    const double snode_eps = 1e-6 * (nn ? T->size[T->root] : 1.0);
    float* sc = (float*)malloc((size_t)(nn ? nn : 1) * 4 * sizeof(float));
    for (int i = 0; i < nn; ++i) {
        sc[4 * i] = (float)T->gcx[i]; sc[4 * i + 1] = (float)T->gcy[i];
        sc[4 * i + 2] = (float)T->gcz[i];
        const double cmax = fmax(fabs(T->gcx[i]), fmax(fabs(T->gcy[i]), fabs(T->gcz[i])));
        sc[4 * i + 3] = (float)(0.5 * T->size[i] + snode_eps + 4.0 * (double)FLT_EPSILON * cmax);
    }
    size_t total = 0;
    for (size_t q = 0; q < nq; ++q) {
        const double hi = hq[q], hi2 = hi * hi;
        int node = nn ? T->root : -1;
        while (node >= 0) {
            const float* sn = sc + 4 * node;
            const double hm = (double)T->hmax[node];
            if (hm <= hi) { node = T->next[node]; continue; }
            double dist = hm + (double)sn[3] + (double)T->xpad[node];
            if (tau > 0) dist += (double)T->vmax[node] * tau;
            const double dx = (double)sn[0] - cx[q];
            if (dx > dist || -dx > dist) { node = T->next[node]; continue; }
            const double dy = (double)sn[1] - cy[q];
            if (dy > dist || -dy > dist) { node = T->next[node]; continue; }
            const double dz = (double)sn[2] - cz[q];
            if (dz > dist || -dz > dist) { node = T->next[node]; continue; }
            dist += 0.7320508075688772 * (double)sn[3];
            if (dx * dx + dy * dy + dz * dz > dist * dist) { node = T->next[node]; continue; }
            if (T->first[node] < 0) {
                for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                    const uint32_t j = T->orderbuf[k];
                    if (j >= n_gas) continue;
                    const double ex = P->x[j] - cx[q], ey = P->y[j] - cy[q], ez = P->z[j] - cz[q];
                    const double r2 = ex * ex + ey * ey + ez * ez;
                    if (r2 >= hi2 && r2 < h[j] * h[j]) {
                        if (found && total < cap) found[total] = j;
                        ++total;
                    }
                }
                node = T->next[node];
            } else {
                node = T->first[node];
            }
        }
        start[q + 1] = total;
    }
    free(sc);
    return total;
}

/* ---------------------------------------------------------------------------------------------
 * The density / smoothing-length solve (hydro.h:24-65 kernels, hydro.cc:150-327), non-periodic,
 * without the neighbour cache or lazy drift
 * ------------------------------------------------------------------------------------------- */

static const double kPiD = 3.14159265358979323846;

/* kernel_norm, ball_vol, kernel_w, kernel_dwdh (hydro.h:29-65), support radius h, dimension dim */
static double knorm(double h, int dim) {
    if (dim == 1) return (4.0 / 3.0) / h;
    if (dim == 2) return 40.0 / (7.0 * kPiD * h * h);
    return 8.0 / (kPiD * h * h * h);
}
static double ball_vol(double h, int dim) {
    if (dim == 1) return 2.0 * h;
    if (dim == 2) return kPiD * h * h;
    return 4.0 * kPiD / 3.0 * h * h * h;
}
double gzo_kernel_w(double r, double h, int dim) {
    const double q = r / h;
    if (q >= 1.0) return 0.0;
    const double norm = knorm(h, dim);
    if (q < 0.5) return norm * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
    const double u = 1.0 - q;
    return norm * 2.0 * u * u * u;
}
double gzo_kernel_dwdh(double r, double h, int dim) {
    const double q = r / h;
    if (q >= 1.0) return 0.0;
    const double norm = knorm(h, dim);
    double f, fp;
    if (q < 0.5) {
        f = 1.0 - 6.0 * q * q + 6.0 * q * q * q;
        fp = -12.0 * q + 18.0 * q * q;
    } else {
        const double u = 1.0 - q;
        f = 2.0 * u * u * u;
        fp = -6.0 * u * u;
    }
    return -((double)dim / h) * norm * f - (q / h) * norm * fp;
}

/* invert_moments (vec.h:43-63): the rank-dim inverse of the symmetric moments (xx yy zz xy yz xz)
 * into b (row-major 3x3, zero outside the live block); 0 when degenerate. 3D is Mat3::invert
 * (math_types/mat3.h:41-67): cofactors times 1 / det, singular when det is 0 or NaN. */
static int invert_moments(const double* e, double* b, int dim) {
    for (int k = 0; k < 9; ++k) b[k] = 0.0;
    if (dim == 1) {
        if (fabs(e[0]) < 1e-300) return 0;
        b[0] = 1.0 / e[0];
        return 1;
    }
    if (dim == 2) {
        const double det = e[0] * e[1] - e[3] * e[3];
        if (fabs(det) < 1e-300) return 0;
        b[0] = e[1] / det; b[1] = -e[3] / det;
        b[3] = -e[3] / det; b[4] = e[0] / det;
        return 1;
    }
    const double m[3][3] = {{e[0], e[3], e[5]}, {e[3], e[1], e[4]}, {e[5], e[4], e[2]}};
    const double d = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (d == 0.0 || d != d) return 0;
    const double id = 1.0 / d;
    b[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    b[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    b[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    b[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    b[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    b[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    b[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    b[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    b[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return 1;
}

/* The particles within radius of c (ngb_search, hydro.cc:37-108, tree aged tau), in traversal order,
 * into *buf (grown as needed); returns the count. sc: the search nodes as gzo_ngb_search packs
 * them. */
static size_t ngb_one(const otree* T, const gzo_particles* P, const float* sc, double c0,
                      double c1, double c2, double radius, double tau, uint32_t** buf,
                      size_t* cap) {
    size_t n = 0;
    const double r2max = radius * radius;
    int node = T->nnodes ? T->root : -1;
    while (node >= 0) {
        const float* sn = sc + 4 * node;
        double dist = radius + (double)sn[3] + (double)T->xpad[node];
        if (tau > 0) dist += (double)T->vmax[node] * tau;
        const double dx = (double)sn[0] - c0;
        if (dx > dist || -dx > dist) { node = T->next[node]; continue; }
        const double dy = (double)sn[1] - c1;
        if (dy > dist || -dy > dist) { node = T->next[node]; continue; }
        const double dz = (double)sn[2] - c2;
        if (dz > dist || -dz > dist) { node = T->next[node]; continue; }
        dist += 0.7320508075688772 * (double)sn[3];
        if (dx * dx + dy * dy + dz * dz > dist * dist) { node = T->next[node]; continue; }
        if (T->first[node] < 0) {
            for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                const uint32_t j = T->orderbuf[k];
                const double ex = P->x[j] - c0, ey = P->y[j] - c1, ez = P->z[j] - c2;
                if (ex * ex + ey * ey + ez * ez < r2max) {
                    if (n == *cap) {
                        *cap = *cap ? 2 * *cap : 256;
                        *buf = (uint32_t*)realloc(*buf, *cap * sizeof(uint32_t));
                    }
                    (*buf)[n++] = j;
                }
            }
            node = T->next[node];
        } else {
            node = T->first[node];
        }
    }
    return n;
}

void gzo_density(void* t, const gzo_particles* P, const uint8_t* type, const uint32_t* targets,
                 size_t nt, double des_ngb, double ngb_tol, const double* h_start, int n_dims,
                 const double* cond_prev, double tau, double* h_out, double* rho_out,
                 int* nngb_out, int* iters_out) {
    const otree* T = (const otree*)t;
    const int nn = T->nnodes;
    const double snode_eps = 1e-6 * (nn ? T->size[T->root] : 1.0);
    float* sc = (float*)malloc((size_t)(nn ? nn : 1) * 4 * sizeof(float));
    for (int i = 0; i < nn; ++i) {
        sc[4 * i] = (float)T->gcx[i]; sc[4 * i + 1] = (float)T->gcy[i];
        sc[4 * i + 2] = (float)T->gcz[i];
        const double cmax = fmax(fabs(T->gcx[i]), fmax(fabs(T->gcy[i]), fabs(T->gcz[i])));
        sc[4 * i + 3] = (float)(0.5 * T->size[i] + snode_eps + 4.0 * (double)FLT_EPSILON * cmax);
    }
    /* default guess: des_ngb particles at the mean density of the bounding box (hydro.cc:170-186) */
    double h_default = 0.0;
    if (!h_start) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (size_t i = 0; i < P->n; ++i) {
            lo[0] = fmin(lo[0], P->x[i]); hi[0] = fmax(hi[0], P->x[i]);
            lo[1] = fmin(lo[1], P->y[i]); hi[1] = fmax(hi[1], P->y[i]);
            lo[2] = fmin(lo[2], P->z[i]); hi[2] = fmax(hi[2], P->z[i]);
        }
        const double bbox_volume = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        const double np = (double)P->n;
        h_default = cbrt(3.0 * des_ngb * bbox_volume / (4.0 * kPiD * np));
        if (n_dims == 1) h_default = 0.5 * des_ngb * (hi[0] - lo[0]) / np;
        if (n_dims == 2) h_default = sqrt(des_ngb * (hi[0] - lo[0]) * (hi[1] - lo[1]) / (kPiD * np));
    }
    uint32_t* nb = NULL;
    size_t cap = 0;
    for (size_t ti = 0; ti < nt; ++ti) {
        const uint32_t tp = targets[ti];
        const double p0 = P->x[tp], p1 = P->y[tp], p2 = P->z[tp];
        double h = h_start ? h_start[ti] : h_default;
        double h_lo = 0.0, h_hi = 0.0;
        int iter = 0, converged = 0, n_inside = 0;
        double rho = 0.0;
        for (; iter < 100; ++iter) {
            const size_t n = ngb_one(T, P, sc, p0, p1, p2, h, tau, &nb, &cap);
            double weight_sum = 0.0, dweight_dh = 0.0;
            double mom[6] = {0, 0, 0, 0, 0, 0}, face[3] = {0, 0, 0};
            rho = 0.0;
            n_inside = 0;
            for (size_t k = 0; k < n; ++k) {
                const uint32_t j = nb[k];
                if (type && type[j] != 0) continue; /* gas neighbours only */
                const double d0 = P->x[j] - p0, d1 = P->y[j] - p1, d2 = P->z[j] - p2;
                const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                const double w = gzo_kernel_w(r, h, n_dims);
                weight_sum += w;
                dweight_dh += gzo_kernel_dwdh(r, h, n_dims);
                rho += P->m[j] * w;
                mom[0] += d0 * d0 * w; mom[1] += d1 * d1 * w; mom[2] += d2 * d2 * w;
                mom[3] += d0 * d1 * w; mom[4] += d1 * d2 * w; mom[5] += d0 * d2 * w;
                face[0] += d0 * w; face[1] += d1 * w; face[2] += d2 * w;
                if (r < h) ++n_inside;
            }
            const double n_eff = ball_vol(h, n_dims) * weight_sum;
            /* the condition-number and face-closure corrections (hydro.cc:220-262) */
            double ncorr = 1.0;
            if (cond_prev) {
                const double cn = cond_prev[tp];
                if (cn > 100.0) ncorr = fmin(sqrt(1.0 + (cn - 100.0) / 1000.0), 2.0);
            }
            double des_eff = des_ngb * ncorr, tol_eff = ngb_tol * ncorr;
            if (weight_sum > 0) {
                double b[9];
                if (invert_moments(mom, b, n_dims)) {
                    const double vol = 1.0 / weight_sum;
                    const double dx_i = sqrt(vol * (mom[0] + mom[1] + mom[2]));
                    double sum_abs = 0.0;
                    for (int c = 0; c < 3; ++c) {
                        const double lk = (b[3 * c] * face[0] + b[3 * c + 1] * face[1] +
                                           b[3 * c + 2] * face[2]) * (2.0 * vol);
                        sum_abs += fabs(lk) / n_dims;
                    }
                    const double denom = 2.0 * n_dims * pow(dx_i, n_dims - 1);
                    if (denom > 0) {
                        const double fce = sum_abs / denom;
                        if (fce > 0.35) ncorr = fmax(ncorr, fmin(fce / 0.35, 2.0));
                        des_eff = des_ngb * ncorr;
                        tol_eff = ngb_tol * ncorr;
                    }
                }
            }
            /* the tolerance widens as the solve iterates (hydro.cc:263-267) */
            if (iter > 1)
                tol_eff = fmin(0.25 * des_eff,
                               tol_eff * exp(0.1 * log(des_eff / (16.0 * tol_eff)) * iter));
            const double residual = n_eff - des_eff;
            if (fabs(residual) < tol_eff) { converged = 1; break; }
            if (residual > 0) h_hi = h;
            else h_lo = h;
            /* Newton on N_eff(h), guarded by the bracket (hydro.cc:271-281) */
            const double dneff_dh =
                ball_vol(h, n_dims) * ((double)n_dims / h * weight_sum + dweight_dh);
            double h_next = (dneff_dh > 0) ? h - residual / dneff_dh : 0.0;
            if (h_next <= h_lo || (h_hi > 0 && h_next >= h_hi) || h_next <= 0) {
                if (h_hi > 0) h_next = (h_lo > 0) ? sqrt(h_lo * h_hi) : 0.5 * h_hi;
                else h_next = h * ((residual > 0) ? 0.7 : 1.6);
            }
            h = h_next;
        }
        if (!converged) { /* rho and the count at the last h (hydro.cc:285-295) */
            const size_t n = ngb_one(T, P, sc, p0, p1, p2, h, tau, &nb, &cap);
            rho = 0.0;
            n_inside = 0;
            for (size_t k = 0; k < n; ++k) {
                const uint32_t j = nb[k];
                if (type && type[j] != 0) continue;
                const double d0 = P->x[j] - p0, d1 = P->y[j] - p1, d2 = P->z[j] - p2;
                const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                if (r < h) ++n_inside;
                rho += P->m[j] * gzo_kernel_w(r, h, n_dims);
            }
        }
        h_out[ti] = h;
        rho_out[ti] = rho;
        nngb_out[ti] = n_inside;
        iters_out[ti] = iter;
    }
    free(nb);
    free(sc);
}

/* ---------------------------------------------------------------------------------------------
 * Walks (tree.cc:528-582, 626-897) and direct summation (tree.cc:946-971)
 * ------------------------------------------------------------------------------------------- */

static inline double soft_of(const gzo_particles* P, size_t i) { return P->soft ? P->soft[i] : 0.0; }

/* Grouped walk (tree.cc:626-897) without tidal/jerk/lazy/sink terms. Groups are runs of targets
 * sharing a leaf when batch <= 0, fixed windows of `batch` otherwise. Opening is decided once per
 * group against its bounding box, largest softening and smallest aold. */
/* grav_tidal_factor (tree.h:86-96): the second kernel derivative, 3/r^5 beyond h. */
static double tidal_factor(double r, double h) {
    const double h_inv = 1.0 / h;
    const double u = r * h_inv;
    if (u >= 1.0) return 3.0 / (r * r * r * r * r);
    const double h_inv5 = h_inv * h_inv * h_inv * h_inv * h_inv;
    double wk;
    if (u < 0.5) wk = 76.8 - 96.0 * u;
    else wk = -0.2 / (u * u * u * u * u) + 48.0 / u - 76.8 + 32.0 * u;
    return wk * h_inv5;
}

/* Tree::node_vel (tree.h:323-331): centre-of-mass velocity corrected for kicks. */
static void node_vel(const otree* T, int no, double* v) {
    const double M = T->mass[no] + (T->dmass ? T->dmass[no] : 0.0);
    if (!T->dp_x || M <= 0) { v[0] = T->vcom_x[no]; v[1] = T->vcom_y[no]; v[2] = T->vcom_z[no]; return; }
    const double m0 = T->mass[no];
    v[0] = (m0 * T->vcom_x[no] + T->dp_x[no]) / M;
    v[1] = (m0 * T->vcom_y[no] + T->dp_y[no]) / M;
    v[2] = (m0 * T->vcom_z[no] + T->dp_z[no]) / M;
}

void gzo_accel_grouped(void* h, const gzo_particles* P, const uint32_t* targets, size_t nt,
                       double theta, double G, int batch, const double* aold, double* ax,
                       double* ay, double* az, long long* counts) {
    gzo_accel_grouped_tj(h, P, targets, nt, theta, G, batch, aold, NULL, NULL, NULL, ax, ay, az,
                         NULL, NULL, counts);
}

/* ... with the tidal tensor (6 x nt, xx yy zz xy yz xz, no G) and the jerk (3 x nt, G included;
 * needs velocities and a tree built with them), accumulated in the same walk (tree.cc:811-871). */
void gzo_accel_grouped_tj(void* h, const gzo_particles* P, const uint32_t* targets, size_t nt,
                          double theta, double G, int batch, const double* aold, const double* vx,
                          const double* vy, const double* vz, double* ax, double* ay,
                          double* az, double* tidal, double* jerk, long long* counts) {
    const otree* T = (const otree*)h;
    const int want_tidal = tidal != NULL;
    const int want_jerk = jerk != NULL && vx != NULL && T->vcom_x != NULL;
    if (jerk && !want_jerk) memset(jerk, 0, 3 * nt * sizeof(double));
    const double theta2 = theta * theta;
    size_t ngroups = 0;
    size_t* gstart = (size_t*)malloc((nt + 2) * sizeof(size_t));
    if (batch <= 0) {
        gstart[ngroups++] = 0;
        for (size_t i = 1; i < nt; ++i) {
            const int li = T->leaf_of[targets[i]], lp = T->leaf_of[targets[i - 1]];
            if (li != lp || li < 0) gstart[ngroups++] = i;
        }
        gstart[ngroups] = nt;
        if (nt == 0) ngroups = 0;
    } else {
        for (size_t i = 0; i < nt; i += (size_t)batch) gstart[ngroups++] = i;
        gstart[ngroups] = nt;
    }
    long long visited = 0, multipole = 0, pairs = 0;

#pragma omp parallel for schedule(dynamic, 1) reduction(+ : visited, multipole, pairs)
    for (long long b = 0; b < (long long)ngroups; ++b) {
        const size_t lo = gstart[b], hi = gstart[b + 1];
        const int nb = (int)(hi - lo);
        double bx0 = 1e300, by0 = 1e300, bz0 = 1e300, bx1 = -1e300, by1 = -1e300, bz1 = -1e300;
        double emax = 0;
        for (size_t i = lo; i < hi; ++i) {
            const uint32_t p = targets[i];
            bx0 = dmin(bx0, P->x[p]); bx1 = dmax(bx1, P->x[p]);
            by0 = dmin(by0, P->y[p]); by1 = dmax(by1, P->y[p]);
            bz0 = dmin(bz0, P->z[p]); bz1 = dmax(bz1, P->z[p]);
            if (P->soft) emax = dmax(emax, P->soft[p]);
        }
        double amin = aold ? 1e300 : 0.0;
        if (aold) for (size_t i = lo; i < hi; ++i) amin = dmin(amin, aold[i]);
        double* buf = (double*)calloc((size_t)nb * 19, sizeof(double));
        double *tx = buf, *ty = buf + nb, *tz = buf + 2 * nb, *te = buf + 3 * nb;
        double *oax = buf + 4 * nb, *oay = buf + 5 * nb, *oaz = buf + 6 * nb;
        double* ott = buf + 7 * nb;                 /* 6 x nb */
        double *tvx = buf + 13 * nb, *tvy = buf + 14 * nb, *tvz = buf + 15 * nb;
        double *ojx = buf + 16 * nb, *ojy = buf + 17 * nb, *ojz = buf + 18 * nb;
        for (int i = 0; i < nb; ++i) {
            const uint32_t p = targets[lo + i];
            tx[i] = P->x[p]; ty[i] = P->y[p]; tz[i] = P->z[p]; te[i] = soft_of(P, p);
            oax[i] = 0; oay[i] = 0; oaz[i] = 0;
            if (want_jerk) { tvx[i] = vx[p]; tvy[i] = vy[p]; tvz[i] = vz[p]; }
        }
        int node = T->root;
        while (node >= 0) {
            ++visited;
            const double wcx = T->wcx[node], wcy = T->wcy[node], wcz = T->wcz[node];
            const double len = (double)T->wlen[node], wsoft = (double)T->wsoft[node];
            const double dx = dmax(0.0, dmax(bx0 - wcx, wcx - bx1));
            const double dy = dmax(0.0, dmax(by0 - wcy, wcy - by1));
            const double dz = dmax(0.0, dmax(bz0 - wcz, wcz - bz1));
            const double rmin2 = dx * dx + dy * dy + dz * dz;
            const int accept = !gzo_open_node(rmin2, len, T->s[node], T->wmass[node], wsoft, emax,
                                              amin, theta2, dx, dy, dz);
            if (accept || T->first[node] < 0) {
                if (!accept) {
                    for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                        const uint32_t q = T->orderbuf[k];
                        const double qx = P->x[q], qy = P->y[q], qz = P->z[q], qm = P->m[q];
                        const double qs = soft_of(P, q);
                        double qvx = 0, qvy = 0, qvz = 0;
                        if (want_jerk) { qvx = vx[q]; qvy = vy[q]; qvz = vz[q]; }
                        for (int i = 0; i < nb; ++i) {
                            if (targets[lo + i] == q) continue;
                            const double ddx = qx - tx[i], ddy = qy - ty[i], ddz = qz - tz[i];
                            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                            if (r2 <= 0) continue;
                            ++pairs;
                            const double r = sqrt(r2);
                            const double fac =
                                qm * gzo_spline_force_over_r(r, dmax(dmax(te[i], qs), 1e-300));
                            oax[i] += ddx * fac; oay[i] += ddy * fac; oaz[i] += ddz * fac;
                            if (want_tidal || want_jerk) { /* tree.cc:811-837 */
                                const double g1 = fac;
                                const double g2 = qm * tidal_factor(r, dmax(dmax(te[i], qs), 1e-300));
                                if (want_tidal) {
                                    ott[i] += -g1 + ddx * ddx * g2;
                                    ott[nb + i] += -g1 + ddy * ddy * g2;
                                    ott[2 * nb + i] += -g1 + ddz * ddz * g2;
                                    ott[3 * nb + i] += ddx * ddy * g2;
                                    ott[4 * nb + i] += ddy * ddz * g2;
                                    ott[5 * nb + i] += ddx * ddz * g2;
                                }
                                if (want_jerk) {
                                    const double dvx = qvx - tvx[i], dvy = qvy - tvy[i],
                                                 dvz = qvz - tvz[i];
                                    const double vdotr = dvx * ddx + dvy * ddy + dvz * ddz;
                                    ojx[i] += g1 * dvx - vdotr * g2 * ddx;
                                    ojy[i] += g1 * dvy - vdotr * g2 * ddy;
                                    ojz[i] += g1 * dvz - vdotr * g2 * ddz;
                                }
                            }
                        }
                    }
                } else {
                    multipole += nb;
                    double nv[3] = {0, 0, 0};
                    if (want_jerk) node_vel(T, node, nv);
                    for (int i = 0; i < nb; ++i) {
                        const double e = dmax(te[i], wsoft);
                        const double ox = wcx - tx[i], oy = wcy - ty[i], oz = wcz - tz[i];
                        const double r2 = ox * ox + oy * oy + oz * oz;
                        if (r2 <= 0) continue;
                        const double f = T->wmass[node] *
                                         gzo_spline_force_over_r(sqrt(r2), dmax(e, 1e-300));
                        oax[i] += ox * f; oay[i] += oy * f; oaz[i] += oz * f;
                        if (want_tidal || want_jerk) { /* tree.cc:845-871 */
                            const double r = sqrt(r2), es = dmax(e, 1e-300);
                            const double g1 = T->wmass[node] * gzo_spline_force_over_r(r, es);
                            const double g2 = T->wmass[node] * tidal_factor(r, es);
                            if (want_tidal) {
                                ott[i] += -g1 + ox * ox * g2;
                                ott[nb + i] += -g1 + oy * oy * g2;
                                ott[2 * nb + i] += -g1 + oz * oz * g2;
                                ott[3 * nb + i] += ox * oy * g2;
                                ott[4 * nb + i] += oy * oz * g2;
                                ott[5 * nb + i] += ox * oz * g2;
                            }
                            if (want_jerk) {
                                const double dvx = nv[0] - tvx[i], dvy = nv[1] - tvy[i],
                                             dvz = nv[2] - tvz[i];
                                const double vdotr = dvx * ox + dvy * oy + dvz * oz;
                                ojx[i] += g1 * dvx - vdotr * g2 * ox;
                                ojy[i] += g1 * dvy - vdotr * g2 * oy;
                                ojz[i] += g1 * dvz - vdotr * g2 * oz;
                            }
                        }
                    }
                }
                node = T->next[node];
            } else {
                node = T->first[node];
            }
        }
        for (int i = 0; i < nb; ++i) {
            ax[lo + i] = G * oax[i]; ay[lo + i] = G * oay[i]; az[lo + i] = G * oaz[i];
            if (want_tidal)
                for (int c = 0; c < 6; ++c) tidal[c * nt + lo + i] = ott[c * nb + i];
            if (want_jerk) {
                jerk[lo + i] = G * ojx[i]; jerk[nt + lo + i] = G * ojy[i];
                jerk[2 * nt + lo + i] = G * ojz[i];
            }
        }
        free(buf);
    }
    free(gstart);
    if (counts) { counts[0] = visited; counts[1] = multipole; counts[2] = pairs; }
}

/* Per-target walk (tree.cc:528-582): a leaf reached is always direct-summed. */
void gzo_accel(void* h, const gzo_particles* P, const uint32_t* targets, size_t nt, double theta,
               double G, const double* aold, double* ax, double* ay, double* az) {
    const otree* T = (const otree*)h;
    const double theta2 = theta * theta;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long t = 0; t < (long long)nt; ++t) {
        const uint32_t p = targets[t];
        const double px = P->x[p], py = P->y[p], pz = P->z[p];
        const double ep = soft_of(P, p);
        const double at = aold ? aold[t] : 0.0;
        double sx = 0, sy = 0, sz = 0;
        int node = T->root;
        while (node >= 0) {
            const double ox = T->wcx[node] - px, oy = T->wcy[node] - py, oz = T->wcz[node] - pz;
            const double r2 = ox * ox + oy * oy + oz * oz;
            if (T->first[node] < 0 ||
                !gzo_open_node(r2, (double)T->wlen[node], T->s[node], T->wmass[node],
                               (double)T->wsoft[node], ep, at, theta2, fabs(ox), fabs(oy),
                               fabs(oz))) {
                if (T->first[node] < 0) {
                    for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                        const uint32_t q = T->orderbuf[k];
                        if (q == p) continue;
                        const double dx = P->x[q] - px, dy = P->y[q] - py, dz = P->z[q] - pz;
                        const double r = sqrt(dx * dx + dy * dy + dz * dz);
                        if (r <= 0) continue;
                        const double fac =
                            P->m[q] * gzo_spline_force_over_r(r, dmax(dmax(ep, soft_of(P, q)), 1e-300));
                        sx += dx * fac; sy += dy * fac; sz += dz * fac;
                    }
                } else {
                    const double e = dmax(ep, (double)T->wsoft[node]);
                    const double r = sqrt(r2);
                    if (r > 0) {
                        const double f = T->wmass[node] * gzo_spline_force_over_r(r, dmax(e, 1e-300));
                        sx += ox * f; sy += oy * f; sz += oz * f;
                    }
                }
                node = T->next[node];
            } else {
                node = T->first[node];
            }
        }
        ax[t] = G * sx; ay[t] = G * sy; az[t] = G * sz;
    }
}

/* The potential walk (tree.cc:899-944): per target, the accel walk's traversal summing the spline
 * potential (tree.h:72-80) of accepted nodes (softening max(target, node)) and of every particle
 * of a reached leaf (softening max of the pair). */
void gzo_potential(void* h, const gzo_particles* P, const uint32_t* targets, size_t nt,
                   double theta, double G, const double* aold, double* phi) {
    const otree* T = (const otree*)h;
    const double theta2 = theta * theta;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long t = 0; t < (long long)nt; ++t) {
        const uint32_t p = targets[t];
        const double px = P->x[p], py = P->y[p], pz = P->z[p];
        const double ep = soft_of(P, p);
        const double at = aold ? aold[t] : 0.0;
        double sum = 0;
        int node = T->root;
        while (node >= 0) {
            const double ox = T->wcx[node] - px, oy = T->wcy[node] - py, oz = T->wcz[node] - pz;
            const double r2 = ox * ox + oy * oy + oz * oz;
            if (T->first[node] < 0 ||
                !gzo_open_node(r2, (double)T->wlen[node], T->s[node], T->wmass[node],
                               (double)T->wsoft[node], ep, at, theta2, fabs(ox), fabs(oy),
                               fabs(oz))) {
                if (T->first[node] < 0) {
                    for (int k = T->plo[node]; k < T->phi[node]; ++k) {
                        const uint32_t q = T->orderbuf[k];
                        if (q == p) continue;
                        const double dx = P->x[q] - px, dy = P->y[q] - py, dz = P->z[q] - pz;
                        const double r = sqrt(dx * dx + dy * dy + dz * dz);
                        if (r <= 0) continue;
                        sum += P->m[q] *
                               gzo_spline_potential(r, dmax(dmax(ep, soft_of(P, q)), 1e-300));
                    }
                } else {
                    const double e = dmax(ep, (double)T->wsoft[node]);
                    const double r = sqrt(r2);
                    if (r > 0) sum += T->wmass[node] * gzo_spline_potential(r, dmax(e, 1e-300));
                }
                node = T->next[node];
            } else {
                node = T->first[node];
            }
        }
        phi[t] = G * sum;
    }
}

/* Exact FP64 sum over every particle with the walk's pair rule (tree.cc:946-971). */
void gzo_accel_brute(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                     double* ax, double* ay, double* az) {
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; ++t) {
        const uint32_t p = targets[t];
        const double px = P->x[p], py = P->y[p], pz = P->z[p];
        const double ep = soft_of(P, p);
        double sx = 0, sy = 0, sz = 0;
        for (size_t q = 0; q < P->n; ++q) {
            if (q == p) continue;
            const double dx = P->x[q] - px, dy = P->y[q] - py, dz = P->z[q] - pz;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= 0) continue;
            const double fac =
                P->m[q] * gzo_spline_force_over_r(r, dmax(dmax(ep, soft_of(P, q)), 1e-300));
            sx += dx * fac; sy += dy * fac; sz += dz * fac;
        }
        ax[t] = G * sx; ay[t] = G * sy; az[t] = G * sz;
    }
}

/* The reference's device direct sum, shmem_cuda_accel_bruteforce (grav_cuda.cu:65-141), in its
 * default float arithmetic (R = float): the separation is differenced in double and narrowed;
 * r, the spline at the larger softening (1e-30 when both are zero), the mass factor and the sums
 * are float, sources in index order, self and r^2 <= 0 skipped. box > 0: nearest periodic image
 * (grav_cuda.cu:104-109). The result is multiplied by G in double (the reference leaves G to its
 * caller). */
static float spline_force_over_r_f(float r, float h) {
    const float h_inv = 1.0f / h;
    const float h_inv3 = h_inv * h_inv * h_inv;
    const float u = r * h_inv;
    if (u >= 1.0f) return 1.0f / (r * r * r);
    if (u < 0.5f) return h_inv3 * (10.666666666666667f + u * u * (32.0f * u - 38.4f));
    return h_inv3 * (21.333333333333332f - 48.0f * u + 38.4f * u * u -
                     10.666666666666667f * u * u * u - 0.06666666666666667f / (u * u * u));
}

static double min_image(double d, double L) {
    const double hb = 0.5 * L;
    if (d > hb) return d - L;
    if (d < -hb) return d + L;
    return d;
}

void gzo_direct_sum_f32(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                        double box, double* ax, double* ay, double* az) {
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; ++t) {
        const uint32_t p = targets[t];
        const double px = P->x[p], py = P->y[p], pz = P->z[p];
        const float ep = (float)soft_of(P, p);
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (size_t q = 0; q < P->n; ++q) {
            if (q == p) continue;
            double dxd = P->x[q] - px, dyd = P->y[q] - py, dzd = P->z[q] - pz;
            if (box > 0) {
                dxd = min_image(dxd, box);
                dyd = min_image(dyd, box);
                dzd = min_image(dzd, box);
            }
            const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= 0.0f) continue;
            const float r = sqrtf(r2);
            const float es = (float)soft_of(P, q);
            float e = ep > es ? ep : es;
            if (!(e > 0.0f)) e = 1e-30f;
            const float fac = (float)P->m[q] * spline_force_over_r_f(r, e);
            sx += fac * dx; sy += fac * dy; sz += fac * dz;
        }
        ax[t] = G * (double)sx; ay[t] = G * (double)sy; az[t] = G * (double)sz;
    }
}

/* accel_brute's FP64 pair rule (tree.cc:946-971) with the nearest periodic image when box > 0;
 * box = 0 is gzo_accel_brute. */
void gzo_direct_sum_box(const gzo_particles* P, const uint32_t* targets, size_t nt, double G,
                        double box, double* ax, double* ay, double* az) {
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; ++t) {
        const uint32_t p = targets[t];
        const double px = P->x[p], py = P->y[p], pz = P->z[p];
        const double ep = soft_of(P, p);
        double sx = 0, sy = 0, sz = 0;
        for (size_t q = 0; q < P->n; ++q) {
            if (q == p) continue;
            double dx = P->x[q] - px, dy = P->y[q] - py, dz = P->z[q] - pz;
            if (box > 0) {
                dx = min_image(dx, box);
                dy = min_image(dy, box);
                dz = min_image(dz, box);
            }
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= 0) continue;
            const double fac =
                P->m[q] * gzo_spline_force_over_r(r, dmax(dmax(ep, soft_of(P, q)), 1e-300));
            sx += dx * fac; sy += dy * fac; sz += dz * fac;
        }
        ax[t] = G * sx; ay[t] = G * sy; az[t] = G * sz;
    }
}

/* audit_tree (tree.cc:973-1119), gravity-only arguments (no h, no box, no lazy drift), over a
 * tree given as plain arrays, so that a device tree downloaded to the host is audited by the same
 * code as the oracle's own. Counts violations and describes the first in msg:
 *   1. threading: from the root, every child chain first[X] .. next[X] ends, no node or particle
 *      is reached twice, parent[] and leaf_of[] agree with the threading, and every particle with
 *      m > 0 is reached;
 *   2. every node's mass equals the sum of its particles' masses (1e-10 relative) and, when
 *      check_com, its centre of mass lies within 1e-6 of its reach (size + 2 (vmax tau + xpad))
 *      of the particles' centre of mass;
 *   3. every ancestor of a particle covers it: softening bound >= its softening, speed bound >=
 *      |v| (when vmax and v are given), and its position within the search node's box (float
 *      geometric centre, half side with the pack's safety epsilon, tree.cc:496-506) widened by
 *      vmax tau + xpad. */
long long gzo_audit_tree(const gzo_tree_view* T, const gzo_particles* P, const double* vx,
                         const double* vy, const double* vz, int check_com, char* msg,
                         size_t msglen) {
    long long bad = 0;
#define AUDIT_FAIL(...)                                                                          \
    do {                                                                                         \
        if (bad++ == 0 && msg && msglen) snprintf(msg, msglen, __VA_ARGS__);                    \
    } while (0)
    const size_t N = P->n, nn = T->nnodes;
    if (msg && msglen) msg[0] = 0;
    if (nn == 0) return 0;
    if (T->root < 0 || (size_t)T->root >= nn) {
        AUDIT_FAIL("root %d out of range %zu", T->root, nn);
        return bad;
    }
    unsigned char* seen_node = calloc(nn, 1);
    unsigned char* seen_part = calloc(N ? N : 1, 1);
    int* stack = malloc((nn + 1) * sizeof(int));
    size_t sp = 0, steps = 0;
    if (T->parent[T->root] != -1) AUDIT_FAIL("root %d has parent %d", T->root, T->parent[T->root]);
    stack[sp++] = T->root;
    while (sp) {
        const int X = stack[--sp];
        if (X < 0 || (size_t)X >= nn) {
            AUDIT_FAIL("node index %d out of range %zu", X, nn);
            continue;
        }
        if (seen_node[X]++) {
            AUDIT_FAIL("node %d reached twice", X);
            continue;
        }
        if (T->first[X] < 0) {
            for (int slot = T->plo[X]; slot < T->phi[X]; ++slot) {
                if (slot < 0 || (size_t)slot >= T->n_live) {
                    AUDIT_FAIL("leaf %d slot %d outside [0, %zu)", X, slot, T->n_live);
                    break;
                }
                const uint32_t q = T->orderbuf[slot];
                if (q >= N) {
                    AUDIT_FAIL("leaf %d holds index %u", X, q);
                    continue;
                }
                if (seen_part[q]++) AUDIT_FAIL("particle %u reached twice, leaf %d", q, X);
                if (T->leaf_of[q] != X)
                    AUDIT_FAIL("leaf_of[%u]=%d, but threading reaches it in leaf %d", q,
                               T->leaf_of[q], X);
            }
            continue;
        }
        for (int c = T->first[X]; c != T->next[X]; c = T->next[c]) {
            if (c < 0 || (size_t)c >= nn || ++steps > 2 * nn) {
                AUDIT_FAIL("child chain of node %d never reaches next[X]=%d", X, T->next[X]);
                break;
            }
            if (T->parent[c] != X) AUDIT_FAIL("parent[%d]=%d, threading says %d", c, T->parent[c], X);
            if (sp < nn + 1) stack[sp++] = c;
        }
    }
    for (size_t q = 0; q < N; ++q)
        if (!seen_part[q] && P->m[q] > 0) {
            AUDIT_FAIL("particle %zu unreachable from the root", q);
            break;
        }
    free(stack);
    free(seen_node);
    free(seen_part);
    if (bad) return bad;

    const double tau = T->tau;
    double* msum = calloc(nn, sizeof(double));
    double* mdx = calloc(3 * nn, sizeof(double));
    for (size_t q = 0; q < N; ++q) {
        if (T->leaf_of[q] < 0 || P->m[q] <= 0) continue;
        for (int no = T->leaf_of[q]; no >= 0; no = T->parent[no]) {
            msum[no] += P->m[q];
            mdx[3 * no + 0] += (P->x[q] - T->cx[no]) * P->m[q];
            mdx[3 * no + 1] += (P->y[q] - T->cy[no]) * P->m[q];
            mdx[3 * no + 2] += (P->z[q] - T->cz[no]) * P->m[q];
        }
    }
    for (size_t no = 0; no < nn; ++no) {
        const double M = T->mass[no];
        const double tol = 1e-10 * fmax(msum[no], fabs(M)) + 1e-300;
        if (fabs(msum[no] - M) > tol) {
            AUDIT_FAIL("node %zu mass %.17g, particles sum to %.17g", no, M, msum[no]);
            break;
        }
        if (check_com && msum[no] > 0) {
            const double ox = mdx[3 * no] / msum[no], oy = mdx[3 * no + 1] / msum[no],
                         oz = mdx[3 * no + 2] / msum[no];
            const double off = sqrt(ox * ox + oy * oy + oz * oz);
            const double vm = T->vmax ? (double)T->vmax[no] : 0.0;
            const double xp = T->xpad ? (double)T->xpad[no] : 0.0;
            const double reach = T->size[no] + 2.0 * (vm * tau + xp);
            if (off > 1e-6 * reach) {
                AUDIT_FAIL("node %zu tracked COM is %.9g off the particles' (node reach %.9g)", no,
                           off, reach);
                break;
            }
        }
    }
    free(msum);
    free(mdx);

    const double snode_eps = 1e-6 * T->size[T->root];
    for (size_t q = 0; q < N && bad < 1000; ++q) {
        if (T->leaf_of[q] < 0 || P->m[q] <= 0) continue;
        const double sq = P->soft ? P->soft[q] : 0.0;
        const double v = (vx && vy && vz) ? sqrt(vx[q] * vx[q] + vy[q] * vy[q] + vz[q] * vz[q]) : 0.0;
        for (int no = T->leaf_of[q]; no >= 0; no = T->parent[no]) {
            if ((double)T->wn_soft[no] * (1.0 + 1e-6) < sq)
                AUDIT_FAIL("node %d soft %.9g below soft=%.9g of particle %zu", no,
                           (double)T->wn_soft[no], sq, q);
            const double vm = T->vmax ? (double)T->vmax[no] : 0.0;
            if (T->vmax && vx && vm * (1.0 + 1e-6) < v)
                AUDIT_FAIL("vmax[%d]=%.9g below |v|=%.9g of particle %zu", no, vm, v, q);
            /* the search node's box (tree.cc:496-506): float centre, half side widened by 1e-6
             * of the root side plus 4 float ulps of the centre's largest coordinate */
            const float fcx = (float)T->gcx[no], fcy = (float)T->gcy[no], fcz = (float)T->gcz[no];
            const double cmax = fmax(fabs(T->gcx[no]), fmax(fabs(T->gcy[no]), fabs(T->gcz[no])));
            const double half =
                (double)(float)(0.5 * T->size[no] + snode_eps + 4.0 * FLT_EPSILON * cmax);
            const double xp = T->xpad ? (double)T->xpad[no] : 0.0;
            const double reach = half + vm * tau + xp + 1e-6 * half;
            const double d[3] = {P->x[q] - (double)fcx, P->y[q] - (double)fcy,
                                 P->z[q] - (double)fcz};
            for (int a = 0; a < 3; ++a)
                if (fabs(d[a]) > reach) {
                    AUDIT_FAIL("particle %zu outside node %d box: offset %.9g > reach %.9g", q, no,
                               fabs(d[a]), reach);
                    break;
                }
        }
    }
#undef AUDIT_FAIL
    return bad;
}
