// Host side of libitomo: error reporting, geometry plans and the C-ABI dispatch.
//
// Per-view constants are computed here, on the host, in the float arithmetic of the
// reference C as its own build compiles it (gcc -O3 -ffast-math: sincosf for the
// trigonometry, reciprocals hoisted out of the loops), or as written (IEEE), so the
// device kernels only multiply, subtract and fuse -- never divide -- in FAST mode.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "itomo_internal.h"

namespace itomo {

static thread_local char g_err[512] = "";

int set_error(int status, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return status;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return ITOMO_OK;
    return set_error(ITOMO_EHIP, "%s: %s", what, hipGetErrorString(e));
}

static const double kPi = 3.141592653589793;  // radon_dd.c:13
static const double kEps = 1.0e-7;            // radon_dd.c:14

static void view_trig(float theta, int arith, float* s, float* c)
{
    if (arith == ITOMO_ARITH_FAST) {
        sincosf(theta, s, c);
    } else {
        *s = (float)std::sin((double)theta);
        *c = (float)std::cos((double)theta);
    }
}

// Branch selection and constants of radon_dd.c:69-283 for one view.
// Whether a pixel span on a sweep line can hold two bin keys.  A span is at most
// 1 + ulp(M) long and two consecutive bin keys are at least r - ulp(M) apart, r = 1/|b|,
// with M bounding every key of the view, so one key per span is certain when
// r - 1 > 2 ulp(M); the flag keeps a factor 2 of margin.
static int dd_multi_cross(const DdView& g, int npix)
{
    const double r = std::fmin((double)g.rinv, 1.0 / std::fabs((double)g.b));
    const double m = (npix + 4.0) * (1.0 + std::fabs((double)g.t) + r);
    return (r - 1.0) <= 4.0 * m * std::ldexp(1.0, -23) ? 1 : 0;
}

static int make_dd_view(float theta, int arith, int row, int npix, DdView* g)
{
    float s, c;
    view_trig(theta, arith, &s, &c);
    std::memset(g, 0, sizeof *g);
    g->row = row;
    g->s = s;
    g->c = c;
    if (std::fabs((double)s) < kEps) {
        g->kind = DD_ROWSUM;
    } else if (std::fabs((double)c) < kEps) {
        g->kind = DD_COLSUM;
    } else if ((double)theta > kPi / 2) {
        if (!(s > 0.0f))
            return set_error(ITOMO_EANGLE,
                             "dd: angle %.9g rad sweeps rows with sin <= 0; the reference "
                             "sweep is undefined there (use angles in [0, pi))",
                             (double)theta);
        g->kind = DD_SWEEP_ROWS;
        g->rinv = 1.0f / s;
        g->ncand = (int)std::floor((double)g->rinv + 0.25) + 2;
        g->t = c * g->rinv;
        g->a = c;
        g->b = s;
        g->est = s;
    } else if ((double)theta < kPi / 2.0 && theta != 0.0f) {
        if (!(c > 0.0f))
            return set_error(ITOMO_EANGLE,
                             "dd: angle %.9g rad sweeps columns with cos <= 0; the reference "
                             "sweep is undefined there (use angles in [0, pi))",
                             (double)theta);
        g->kind = DD_SWEEP_COLS;
        g->rinv = 1.0f / c;
        g->ncand = (int)std::floor((double)g->rinv + 0.25) + 2;
        g->t = s * g->rinv;
        g->a = s;
        g->b = c;
        g->est = c;
    } else {
        g->kind = DD_NONE;
    }
    if (g->kind == DD_SWEEP_ROWS || g->kind == DD_SWEEP_COLS) g->multi = dd_multi_cross(*g, npix);
    return ITOMO_OK;
}

static int make_rd_view(float theta, int arith, RdView* g)
{
    float s, c;
    view_trig(theta, arith, &s, &c);
    std::memset(g, 0, sizeof *g);
    g->theta = theta;
    g->s = s;
    g->c = c;
    g->lt_half_pi = ((double)theta < kPi / 2) ? 1 : 0;
    if (std::fabs((double)s) < kEps) {
        g->kind = RD_COLSUM;
    } else if (std::fabs((double)c) < kEps) {
        g->kind = RD_ROWSUM;
    } else {
        g->kind = RD_RAYS;
        g->rc = 1.0f / c;
        g->rs = 1.0f / s;
    }
    return ITOMO_OK;
}

static void make_pd_view(float theta, int arith, PdView* g)
{
    float s, c;
    view_trig(theta, arith, &s, &c);
    std::memset(g, 0, sizeof *g);
    g->s = s;
    g->c = c;
    g->rc = 1.0f / c;
    g->flags = ((double)theta < kPi / 2 ? kPdLt : 0) | ((double)theta > kPi / 2 ? kPdGt : 0) |
               (theta == 0.0f ? kPdZero : 0) | ((double)theta == kPi / 2 ? kPdHalf : 0);
    g->cabs = std::fabs((double)c);
    g->sdir = c < 0.0f ? -(double)s : (double)s;
    g->rsdir = g->sdir != 0.0 ? 1.0 / g->sdir : 0.0;
    g->dshift = 0.25 * std::fabs(g->cabs - g->sdir);
}

static const double kSsEps = 1e-5;                   // radon_ss.c:11
static const double kSsS45 = 0.70710678118654757;    // radon_ss.c:12

static void make_ss_view(float theta, int arith, SsView* g)
{
    float s, c;
    view_trig(theta, arith, &s, &c);
    std::memset(g, 0, sizeof *g);
    g->s = s;
    g->c = c;
    g->ybranch = kSsS45 < std::fabs((double)s) ? 0 : 1;
    g->a = g->ybranch ? s : c;
    g->b = g->ybranch ? c : s;
    g->rb = 1.0f / g->b;
    g->rab = 1.0 / std::fabs((double)g->b);
}

static int ss_round_host(float v, int arith)
{
    if (arith == ITOMO_ARITH_FAST) {  // the fast build's float expansion of round()
        const float r = v + std::copysign(0.49999997f, v);
        return (int)std::trunc(r);
    }
    return (int)std::round((double)v);
}

// Middle two of {p, q, lo, hi} (lo <= hi), ascending: the reference sorts the four
// values and keeps elements 1 and 2 (radon_ss.c:76-84).
static void ss_middle_two(float p, float q, float lo, float hi, float* m, float* n)
{
    const float mn = p < q ? p : q, mx = p < q ? q : p;
    const float m1 = mn > lo ? mn : lo, m2 = mx < hi ? mx : hi;
    *m = m1 < m2 ? m1 : m2;
    *n = m1 < m2 ? m2 : m1;
}

// Stepped range of every bin of one view (ray_limits, radon_ss.c:43-88, and the rounded
// loop bounds of radon_ss.c:121,159), clipped to the image [-nh, npix-nh).
static void ss_view_ranges(const SsView& g, int arith, int npix, int* out)
{
#pragma clang fp contract(off)
    const int nh = (int)(npix * 0.5);
    const float s = g.s, c = g.c;
    const float nhf = (float)nh, nh1 = (float)(nh - 1), mnh = (float)(-nh);
    for (int k = 0; k < npix; ++k) {
        int u0 = 0, u1 = 0;
        if (k < 2 * nh) {
            const int t = k - nh;
            const float tf = (float)t;
            float lim[4];
            if (std::fabs((double)s) < kSsEps) {
                lim[0] = 0.0f; lim[1] = 0.0f; lim[2] = mnh; lim[3] = nh1;
            } else if (std::fabs((double)c) < kSsEps) {
                lim[0] = mnh; lim[1] = nh1; lim[2] = 0.0f; lim[3] = 0.0f;
            } else {
                float x1, x2, y1, y2;
                if (arith == ITOMO_ARITH_FAST) {
                    x1 = std::fma(nhf, s, tf) / c;
                    x2 = std::fma(-s, nh1, tf) / c;
                    y1 = std::fma(nhf, c, tf) / s;
                    y2 = std::fma(-c, nh1, tf) / s;
                } else {
                    x1 = tf / c + (nhf * s) / c;
                    x2 = tf / c - (nh1 * s) / c;
                    y1 = tf / s + (nhf * c) / s;
                    y2 = tf / s - (nh1 * c) / s;
                }
                ss_middle_two(x1, x2, mnh, nh1, &lim[0], &lim[1]);
                ss_middle_two(y1, y2, mnh, nh1, &lim[2], &lim[3]);
            }
            u0 = ss_round_host(g.ybranch ? lim[2] : lim[0], arith);
            u1 = ss_round_host(g.ybranch ? lim[3] : lim[1], arith);
            u0 = u0 < -nh ? -nh : u0;
            u1 = u1 > npix - nh ? npix - nh : u1;
            if (u1 < u0) u1 = u0;
        }
        out[2 * k] = u0;
        out[2 * k + 1] = u1;
    }
}

int plan_scratch(const Plan& p, size_t bytes, void** ptr, hipStream_t st)
{
    if (bytes > p.scratch_bytes) {
        // queued work may still read the old buffer
        int rc = check_hip(hipStreamSynchronize(st), "plan scratch: sync");
        if (rc) return rc;
        if (p.scratch) {
            rc = check_hip(hipFree(p.scratch), "plan scratch: hipFree");
            p.scratch = nullptr;
            p.scratch_bytes = 0;
            if (rc) return rc;
        }
        const size_t want = (bytes + 4095) & ~(size_t)4095;
        rc = check_hip(hipMalloc(&p.scratch, want), "plan scratch: hipMalloc");
        if (rc) return rc;
        p.scratch_bytes = want;
    }
    *ptr = p.scratch;
    return ITOMO_OK;
}

namespace {

struct ScratchLayout {
    size_t o_pf, o_pb, o_img, total;
};

ScratchLayout scratch_layout(const Plan& p, int nslices)
{
    const size_t ns = (size_t)nslices, n2 = (size_t)p.npix * p.npix;
    const size_t nxt = n2 * ns;
    const size_t nf = fwd_blocks_per_slice(p) * 3 * ns;
    // back partials also serve the elementwise SPS update (one slot per kEpiChunk pixels)
    const size_t nbb = back_blocks_per_slice(p), nbe = (n2 + kEpiChunk - 1) / kEpiChunk;
    const size_t nb = (nbb > nbe ? nbb : nbe) * 3 * ns;
    const size_t ni = p.projector == ITOMO_PROJ_RD ? n2 * ns : 0;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    ScratchLayout l;
    l.o_pf = up(nxt * sizeof(float));
    l.o_pb = l.o_pf + up(nf * sizeof(double));
    l.o_img = l.o_pb + up(nb * sizeof(double));
    l.total = l.o_img + up(ni * sizeof(float));
    return l;
}

thread_local ScratchRegion t_region = {nullptr, 0};

}  // namespace

bool scratch_region_active() { return t_region.base != nullptr; }

ScratchScope::ScratchScope(ScratchRegion r) : prev_(t_region) { t_region = r; }
ScratchScope::~ScratchScope() { t_region = prev_; }

size_t scratch_total(const Plan& p, int nslices) { return scratch_layout(p, nslices).total; }

int get_scratch(const Plan& p, int nslices, Scratch* s, hipStream_t st)
{
    const ScratchLayout l = scratch_layout(p, nslices);
    void* base = nullptr;
    if (t_region.base) {
        if (l.total > t_region.bytes)
            return set_error(ITOMO_EINVAL, "scratch: %zu bytes needed for %d slice(s) of %dx%d, "
                                           "the supplied region holds %zu",
                             l.total, nslices, p.npix, p.npix, t_region.bytes);
        base = t_region.base;
    } else {
        int rc = plan_scratch(p, l.total, &base, st);
        if (rc) return rc;
    }
    char* b = (char*)base;
    s->xT = (float*)b;
    s->pf = (double*)(b + l.o_pf);
    s->pb = (double*)(b + l.o_pb);
    s->img = (float*)(b + l.o_img);
    return ITOMO_OK;
}

}  // namespace itomo

using namespace itomo;

extern "C" {

const char* itomo_version(void) { return "itomo 0.1.0 (gfx950)"; }
const char* itomo_last_error(void) { return g_err; }

static int plan_create(itomo_plan** out, int projector, int npix, const float* angles,
                       int nang, int method, int arith)
{
    if (!out) return set_error(ITOMO_EINVAL, "plan_create: null output pointer");
    *out = nullptr;
    if (projector != ITOMO_PROJ_DD && projector != ITOMO_PROJ_RD &&
        projector != ITOMO_PROJ_PD && projector != ITOMO_PROJ_SS)
        return set_error(ITOMO_EINVAL, "plan_create: unknown projector %d", projector);
    if (method != 0 && method != 1)
        return set_error(ITOMO_EINVAL, "plan_create: method %d is neither 0 (nearest) nor 1 "
                                       "(linear)", method);
    if (arith != ITOMO_ARITH_FAST && arith != ITOMO_ARITH_IEEE)
        return set_error(ITOMO_EINVAL, "plan_create: unknown arithmetic %d", arith);
    // short indices in radon_dd.c:19-20 bound npix to 32766; keep well inside int math
    if (npix < 1 || npix > 32766)
        return set_error(ITOMO_EINVAL, "plan_create: npix %d outside [1, 32766]", npix);
    if (nang < 1 || !angles)
        return set_error(ITOMO_EINVAL, "plan_create: need nang >= 1 host angles");
    for (int v = 0; v < nang; ++v)
        if (!std::isfinite(angles[v]))
            return set_error(ITOMO_EANGLE, "plan_create: angle %d is not finite", v);

    itomo_plan* p = new (std::nothrow) itomo_plan();
    if (!p) return set_error(ITOMO_ENOMEM, "plan_create: out of host memory");
    p->projector = projector;
    p->npix = npix;
    p->nang = nang;
    p->arith = arith;
    p->method = method;
    size_t bytes = 0;
    int st = ITOMO_OK;
    if (projector == ITOMO_PROJ_DD) {
        bytes = sizeof(DdView) * (size_t)nang;
        DdView* h = (DdView*)std::malloc(bytes);
        if (!h) { delete p; return set_error(ITOMO_ENOMEM, "plan_create: out of host memory"); }
        for (int v = 0; v < nang && st == ITOMO_OK; ++v)
            st = make_dd_view(angles[v], arith, nang - 1 - v, npix, &h[v]);
        p->h_geom = h;
    } else if (projector == ITOMO_PROJ_SS) {
        bytes = sizeof(SsView) * (size_t)nang;
        SsView* h = (SsView*)std::malloc(bytes);
        if (!h) { delete p; return set_error(ITOMO_ENOMEM, "plan_create: out of host memory"); }
        for (int v = 0; v < nang; ++v) make_ss_view(angles[v], arith, &h[v]);
        p->h_geom = h;
    } else if (projector == ITOMO_PROJ_PD) {
        bytes = sizeof(PdView) * (size_t)nang;
        PdView* h = (PdView*)std::malloc(bytes);
        if (!h) { delete p; return set_error(ITOMO_ENOMEM, "plan_create: out of host memory"); }
        for (int v = 0; v < nang; ++v) make_pd_view(angles[v], arith, &h[v]);
        p->h_geom = h;
    } else {
        bytes = sizeof(RdView) * (size_t)nang;
        RdView* h = (RdView*)std::malloc(bytes);
        if (!h) { delete p; return set_error(ITOMO_ENOMEM, "plan_create: out of host memory"); }
        for (int v = 0; v < nang && st == ITOMO_OK; ++v)
            st = make_rd_view(angles[v], arith, &h[v]);
        p->h_geom = h;
    }
    if (st != ITOMO_OK) {
        std::free(p->h_geom);
        delete p;
        return st;
    }
    st = check_hip(hipMalloc(&p->d_geom, bytes), "plan_create: hipMalloc");
    if (st == ITOMO_OK)
        st = check_hip(hipMemcpy(p->d_geom, p->h_geom, bytes, hipMemcpyHostToDevice),
                       "plan_create: hipMemcpy");
    if (st == ITOMO_OK && projector == ITOMO_PROJ_DD) st = dd_build_line_ends(*p);
    if (st == ITOMO_OK && projector == ITOMO_PROJ_DD) st = dd_build_view_lists(*p);
    if (st == ITOMO_OK && projector == ITOMO_PROJ_DD && p->n_staged > 0 && p->n_gather > 0) {
        st = check_hip(hipStreamCreateWithFlags(&p->aux_stream, hipStreamNonBlocking),
                       "plan_create: stream");
        if (st == ITOMO_OK)
            st = check_hip(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming),
                           "plan_create: event");
        if (st == ITOMO_OK)
            st = check_hip(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming),
                           "plan_create: event");
    }
    if (st == ITOMO_OK && projector == ITOMO_PROJ_SS) {
        const size_t n = (size_t)nang * npix * 2;
        std::vector<int> tab;
        try {
            tab.resize(n);
        } catch (const std::bad_alloc&) {
            st = set_error(ITOMO_ENOMEM, "plan_create: out of host memory");
        }
        if (st == ITOMO_OK) {
            const SsView* g = (const SsView*)p->h_geom;
            for (int v = 0; v < nang; ++v)
                ss_view_ranges(g[v], arith, npix, tab.data() + (size_t)v * npix * 2);
            st = check_hip(hipMalloc(&p->d_range, n * sizeof(int)), "plan_create: hipMalloc");
        }
        if (st == ITOMO_OK)
            st = check_hip(hipMemcpy(p->d_range, tab.data(), n * sizeof(int),
                                     hipMemcpyHostToDevice),
                           "plan_create: hipMemcpy");
    }
    if (st != ITOMO_OK) {
        if (p->ev_join) (void)hipEventDestroy(p->ev_join);
        if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
        if (p->aux_stream) (void)hipStreamDestroy(p->aux_stream);
        if (p->d_vlist) (void)hipFree(p->d_vlist);
        if (p->d_range) (void)hipFree(p->d_range);
        if (p->d_lstar) (void)hipFree(p->d_lstar);
        if (p->d_geom) (void)hipFree(p->d_geom);
        std::free(p->h_geom);
        delete p;
        return st;
    }
    *out = p;
    return ITOMO_OK;
}

int itomo_plan_create(itomo_plan** out, int projector, int npix, const float* angles,
                      int nang, int arith)
{
    return plan_create(out, projector, npix, angles, nang, 1, arith);
}

int itomo_plan_create_pd(itomo_plan** out, int npix, const float* angles, int nang,
                         int method, int arith)
{
    return plan_create(out, ITOMO_PROJ_PD, npix, angles, nang, method, arith);
}

int itomo_plan_create_ss(itomo_plan** out, int npix, const float* angles, int nang,
                         int method, int arith)
{
    return plan_create(out, ITOMO_PROJ_SS, npix, angles, nang, method, arith);
}

int itomo_plan_destroy(itomo_plan* p)
{
    if (!p) return ITOMO_OK;
    int st = ITOMO_OK;
    if (p->d_geom) st = check_hip(hipFree(p->d_geom), "plan_destroy: hipFree");
    if (p->d_lstar) {
        int st2 = check_hip(hipFree(p->d_lstar), "plan_destroy: hipFree");
        if (!st) st = st2;
    }
    if (p->d_range) {
        int st2 = check_hip(hipFree(p->d_range), "plan_destroy: hipFree");
        if (!st) st = st2;
    }
    if (p->d_vlist) {
        int st2 = check_hip(hipFree(p->d_vlist), "plan_destroy: hipFree");
        if (!st) st = st2;
    }
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->aux_stream) {
        int st2 = check_hip(hipStreamDestroy(p->aux_stream), "plan_destroy: stream");
        if (!st) st = st2;
    }
    if (p->scratch) {
        int st2 = check_hip(hipFree(p->scratch), "plan_destroy: hipFree scratch");
        if (!st) st = st2;
    }
    rd_matrix_free(*p);
    std::free(p->h_geom);
    delete p;
    return st;
}

int itomo_plan_trig(const itomo_plan* p, float* s_host, float* c_host)
{
    if (!p || !s_host || !c_host) return set_error(ITOMO_EINVAL, "plan_trig: null pointer");
    if (p->projector == ITOMO_PROJ_RD) {
        const RdView* g = (const RdView*)p->h_geom;
        for (int v = 0; v < p->nang; ++v) { s_host[v] = g[v].s; c_host[v] = g[v].c; }
    } else if (p->projector == ITOMO_PROJ_PD) {
        const PdView* g = (const PdView*)p->h_geom;
        for (int v = 0; v < p->nang; ++v) { s_host[v] = g[v].s; c_host[v] = g[v].c; }
    } else if (p->projector == ITOMO_PROJ_SS) {
        const SsView* g = (const SsView*)p->h_geom;
        for (int v = 0; v < p->nang; ++v) { s_host[v] = g[v].s; c_host[v] = g[v].c; }
    } else {
        const DdView* g = (const DdView*)p->h_geom;
        for (int v = 0; v < p->nang; ++v) { s_host[v] = g[v].s; c_host[v] = g[v].c; }
    }
    return ITOMO_OK;
}

int itomo_plan_set_shapes(itomo_plan* p, int fwd_variant, int back_variant)
{
    if (!p) return set_error(ITOMO_EINVAL, "plan_set_shapes: null plan");
    const int nf = p->projector == ITOMO_PROJ_DD ? kNumFwdVariants : 1;
    const int nb = p->projector == ITOMO_PROJ_DD ? kNumBackVariants : 1;
    if (fwd_variant < 0 || fwd_variant >= nf || back_variant < 0 || back_variant >= nb)
        return set_error(ITOMO_EINVAL, "plan_set_shapes: variants (%d, %d) outside [0, %d) x "
                                       "[0, %d)", fwd_variant, back_variant, nf, nb);
    p->fwd_var = fwd_variant;
    p->back_var = back_variant;
    return ITOMO_OK;
}

int itomo_plan_get_shapes(const itomo_plan* p, int* fwd_variant, int* back_variant)
{
    if (!p || !fwd_variant || !back_variant)
        return set_error(ITOMO_EINVAL, "plan_get_shapes: null pointer");
    *fwd_variant = p->fwd_var;
    *back_variant = p->back_var;
    return ITOMO_OK;
}

static int check_io(const itomo_plan* p, const void* a, const void* b, int nslices,
                    const char* who)
{
    if (!p) return set_error(ITOMO_EINVAL, "%s: null plan", who);
    if (!a || !b) return set_error(ITOMO_EINVAL, "%s: null array", who);
    if (nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "%s: nslices %d outside [1, 65535]", who, nslices);
    return ITOMO_OK;
}

int itomo_forward(const itomo_plan* p, const float* image, float* sino, int nslices,
                  int accumulate, void* stream)
{
    int st = check_io(p, image, sino, nslices, "forward");
    if (st) return st;
    const int64_t is = (int64_t)p->npix * p->npix, ss = (int64_t)p->nang * p->npix;
    hipStream_t s = (hipStream_t)stream;
    if (p->projector == ITOMO_PROJ_DD)
        return dd_forward(*p, image, sino, nslices, is, ss, accumulate, s);
    if (p->projector == ITOMO_PROJ_PD)
        return pd_forward(*p, image, sino, nslices, is, ss, accumulate, s);
    if (p->projector == ITOMO_PROJ_SS)
        return ss_forward(*p, image, sino, nslices, is, ss, accumulate, s);
    return rd_forward(*p, image, sino, nslices, is, ss, accumulate, s);
}

int itomo_back(const itomo_plan* p, const float* sino, float* image, int nslices,
               int accumulate, void* stream)
{
    int st = check_io(p, sino, image, nslices, "back");
    if (st) return st;
    const int64_t is = (int64_t)p->npix * p->npix, ss = (int64_t)p->nang * p->npix;
    hipStream_t s = (hipStream_t)stream;
    if (p->projector == ITOMO_PROJ_DD)
        return dd_back(*p, sino, image, nslices, ss, is, accumulate, s);
    if (p->projector == ITOMO_PROJ_PD)
        return pd_back(*p, sino, image, nslices, ss, is, accumulate, s);
    if (p->projector == ITOMO_PROJ_SS)
        return ss_back(*p, sino, image, nslices, ss, is, accumulate, s);
    return rd_back(*p, sino, image, nslices, ss, is, accumulate, s);
}

static int one_shot(int projector, float* image, int npix, const float* angles, int nang,
                    int oper, int method, float* sino, void* stream)
{
    if (oper != 0 && oper != 1)
        return set_error(ITOMO_EINVAL, "radon: oper must be 0 (forward) or 1 (back)");
    itomo_plan* p = nullptr;
    int st = plan_create(&p, projector, npix, angles, nang, method, ITOMO_ARITH_FAST);
    if (st) return st;
    st = oper == 0 ? itomo_forward(p, image, sino, 1, 1, stream)
                   : itomo_back(p, sino, image, 1, 1, stream);
    // the plan's device table must outlive the queued kernel
    int st2 = check_hip(hipStreamSynchronize((hipStream_t)stream), "radon: sync");
    int st3 = itomo_plan_destroy(p);
    return st ? st : (st2 ? st2 : st3);
}

int itomo_radon_dd(float* image, int npix, const float* angles, int nang, int oper,
                   float* sino, void* stream)
{
    return one_shot(ITOMO_PROJ_DD, image, npix, angles, nang, oper, 1, sino, stream);
}

int itomo_radon_rd(float* image, int npix, const float* angles, int nang, int oper,
                   float* sino, void* stream)
{
    return one_shot(ITOMO_PROJ_RD, image, npix, angles, nang, oper, 1, sino, stream);
}

int itomo_radon_pd(float* image, int npix, const float* angles, int nang, int oper, int method,
                   float* sino, void* stream)
{
    return one_shot(ITOMO_PROJ_PD, image, npix, angles, nang, oper, method, sino, stream);
}

int itomo_radon_ss(float* image, int npix, const float* angles, int nang, int oper, int method,
                   float* sino, void* stream)
{
    return one_shot(ITOMO_PROJ_SS, image, npix, angles, nang, oper, method, sino, stream);
}

}  // extern "C"
