// Internal declarations shared by the HIP translation units of libitomo.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "itomo.h"

namespace itomo {

// Per-view geometry kinds of the distance-driven projector (radon_dd.c:75-283).
enum DdKind : int {
    DD_NONE = 0,   // no branch taken (theta == 0 with |sin| >= 1e-7, or NaN)
    DD_ROWSUM = 1, // |sin| < 1e-7: sino[row][i] += sum_j img[i][j]
    DD_COLSUM = 2, // |cos| < 1e-7: sino[row][i] += sum_j img[j][i]
    DD_SWEEP_ROWS = 3, // theta > pi/2: sweep along image rows
    DD_SWEEP_COLS = 4, // 0 < theta < pi/2: sweep along image columns
};

// Per-view geometry kinds of the ray-driven projector (radon_rd.c:96-187).
enum RdKind : int {
    RD_COLSUM = 1,  // |sin| < 1e-7: sino[v][N-1-i] += sum_j img[j][i], i >= 1
    RD_ROWSUM = 2,  // |cos| < 1e-7: sino[v][i] += sum_j img[i][j]
    RD_RAYS = 3,    // Siddon traversal of 5 sub-rays per bin
};

// Constants of one distance-driven view, computed on the host in float exactly as
// the reference's compiled C computes them.
//   line coordinate  u_i = (i - nh) + 0.5
//   pixel boundaries P_j = (j - nh) - K_i      (j = 0..N)
//   bin boundaries   D_d = (d - nh) * rinv     (FAST) or (d - nh) / b (IEEE)
//   K_i = u_i * t (FAST) or (u_i * a) / b (IEEE)
struct DdView {
    int kind;
    int row;     // output sinogram row (nang-1-v)
    float rinv;  // 1/sin (rows) or 1/cos (columns)
    float t;     // cos*rinv (rows) or sin*rinv (columns)
    float a, b;  // IEEE form: (cos, sin) for rows, (sin, cos) for columns
    float est;   // b, used only to estimate a bin index (result is corrected exactly)
    float s, c;  // sin, cos of the view as the reference computes them
    int ncand;   // pixels that can overlap one bin on one line: floor(rinv + 0.25) + 2
    int multi;   // a pixel's span can hold two bin keys (bin width within rounding of 1)
};

struct RdView {
    int kind;
    int lt_half_pi;  // theta < pi/2 (double compare, radon_rd.c:125-128)
    float theta, s, c;
    float rc, rs;    // 1/c, 1/s (FAST form)
    float pad;
};

// Per-view constants of the pixel-driven projector (radon_pd.c:29-55).
enum PdFlags : int {
    kPdLt = 1,    // (double)theta < pi/2
    kPdGt = 2,    // (double)theta > pi/2
    kPdZero = 4,  // theta == 0
    kPdHalf = 8,  // (double)theta == pi/2 (never true for a float angle; kept as written)
};

struct PdView {
    float s, c;   // sin, cos of the view as the reference computes them
    float rc;     // 1/c (FAST form of x*s/c)
    int flags;    // PdFlags
    // Candidate ranges only (never in the arithmetic of a result): the sample distance
    // is t ~ cabs*y - sdir*x; rsdir = 1/sdir, or 0 when sdir == 0; the two sample
    // points of a pixel sit at +-dshift = +-|cabs - sdir|/4 from its centre's t.
    double cabs, sdir, rsdir, dshift;
};

// Per-view constants of the slant-stacking projector (radon_ss.c:103-128).  Views with
// |s| <= sqrt(1/2) step image rows u = y and sample columns q = (t - u s)/c; the others
// step columns u = x and sample rows q = (t - u c)/s: q = (t - u a)/b.
struct SsView {
    float s, c;   // sin, cos as the reference computes them
    float a, b;   // (s, c) or (c, s)
    float rb;     // 1/b (FAST form of the division by b)
    int ybranch;  // 1: step rows (|s| <= sqrt(1/2))
    double rab;   // 1/|b| in double (radon_ss.c:128 `1.0/fabs(c)`)
};

struct Plan {
    int projector;   // ITOMO_PROJ_*
    int npix;
    int nang;
    int arith;       // ITOMO_ARITH_*
    void* d_geom;    // device array of DdView / RdView
    void* h_geom;    // host copy
    float* d_lstar;  // distance-driven: [view][line] start of the dropped segment
    int* d_range;    // slant-stacking: [view][bin] stepped range [u0, u1) (int2), clipped
                     // to the image
    int* d_vlist;    // distance-driven: views of the main forward launch, then the wide ones
    int n_staged, n_gather;  // (more than 8 candidate pixels per bin and line)
    // Distance-driven plans with both kinds of views: the wide views' forward launch runs
    // on this second stream (default priority), forked from and joined back into the
    // caller's stream by the two events, beside the main launch.
    hipStream_t aux_stream;
    hipEvent_t ev_fork, ev_join;
    int fwd_var;     // distance-driven launch shapes (kFwdBins / kBackTiles index)
    int back_var;
    int method;      // pixel-driven: 0 nearest bin, 1 linear split (radon_pd.c:57-85)
    // Device scratch owned by the plan (FP64 partial sums, ray-driven back-projection
    // image).  Grown on first use; a plan must not be used from two streams at once.
    mutable void* scratch;
    mutable size_t scratch_bytes;
    // Ray-driven plans: the projector's exact system matrix cached in HBM (rd.hip), built
    // on the first iteration call when it fits kRdMatrixMaxBytes.  Forward rows (one per
    // bin) and back columns (one per pixel) in chunks of 64 outputs, entry j of lane k of
    // a chunk at off[chunk] + 64 j + k, cnt[out] = the output's own number of entries; an
    // entry is {input index, length bits}.
    mutable int rd_mat_state;   // 0: not tried, 1: built, -1: declined
    mutable int rd_mat_pinned;  // 1 once a stream capture has recorded a launch using it
    mutable int2* rdf_ent;
    mutable int64_t* rdf_off;
    mutable int* rdf_cnt;
    mutable int2* rdb_ent;
    mutable int64_t* rdb_off;
    mutable int* rdb_cnt;
    mutable size_t rd_mat_bytes;
};

// Largest ray-driven system matrix a plan caches (forward and back entries together).
constexpr size_t kRdMatrixMaxBytes = (size_t)8 << 30;
// Builds the ray-driven system matrix of p if not tried yet and allowed (no caller-owned
// scratch region active); synchronous.  Returns ITOMO_OK whether built or declined.
int rd_matrix_ensure(const Plan& p, hipStream_t st);
void rd_matrix_free(const Plan& p);
// out = the transposes of nslices npix x npix images (npix^2 apart); stream-ordered.
int launch_transpose(const float* in, float* out, int npix, int nslices, hipStream_t st);
// Marks the matrix as held by a graph when st is capturing.
void rd_matrix_note_use(const Plan& p, hipStream_t st);
// Whether a caller-owned scratch region is active on this thread (no allocation then).
bool scratch_region_active();

// Returns plan scratch of at least `bytes` bytes (allocating on first use / growth).
int plan_scratch(const Plan& p, size_t bytes, void** ptr, hipStream_t st);

// Launch geometry shared by the launchers and the scratch layout.
// Distance-driven launch shapes (every shape gives bit-identical projections):
//   forward: kFwdBins[v] bins of one view per block, one bin per thread;
//   back:    kBackTiles[v] = {T, P}: T x T pixel tile per block, P x P per thread.
constexpr int kNumFwdVariants = 3;
constexpr int kFwdBins[kNumFwdVariants] = {256, 128, 64};
constexpr int kNumBackVariants = 3;
constexpr int kBackTiles[kNumBackVariants][2] = {{64, 4}, {32, 2}, {32, 4}};
constexpr int kDdFwdBlock = 256;            // default forward shape
constexpr int kDdBackTile = 64;             // default back tile (also the batch's)
constexpr int kRdBlock = 64;
constexpr int kEpiChunk = 256 * 8;          // pixels per elementwise-epilogue block

// FP64 partial-sum slots of one forward slice: one per 64-bin chunk of a view (the
// forward reduces per wave; a wave holds the chunk's 64 bins whatever the block width),
// slot view * ceil(npix / 64) + chunk.  Every slot is written by exactly one wave, also
// when launches of different block widths share the table.
inline size_t dd_fwd_slots(int npix, int nang)
{
    return (size_t)((npix + 63) / 64) * nang;
}

inline size_t dd_back_blocks(int npix, int tile)
{
    const size_t t = (size_t)((npix + tile - 1) / tile);
    return t * t;
}

inline size_t fwd_blocks_per_slice(const Plan& p)
{
    if (p.projector == ITOMO_PROJ_DD) return dd_fwd_slots(p.npix, p.nang);
    if (p.projector == ITOMO_PROJ_PD || p.projector == ITOMO_PROJ_SS)
        // partials come from the elementwise epilogue pass
        return ((size_t)p.nang * p.npix + kEpiChunk - 1) / kEpiChunk;
    return (size_t)((p.npix + kRdBlock - 1) / kRdBlock) * p.nang;
}

// Pixel-driven and slant-stacking back projection: 64 x 4 pixels per block.
inline size_t pd_back_blocks(int npix)
{
    return (size_t)((npix + 63) / 64) * (size_t)((npix + 3) / 4);
}

inline size_t back_blocks_per_slice(const Plan& p)
{
    const size_t n2 = (size_t)p.npix * p.npix;
    if (p.projector == ITOMO_PROJ_DD) return dd_back_blocks(p.npix, kBackTiles[p.back_var][0]);
    if (p.projector == ITOMO_PROJ_PD || p.projector == ITOMO_PROJ_SS)
        return pd_back_blocks(p.npix);
    return (n2 + kEpiChunk - 1) / kEpiChunk;
}

// Per-call device scratch carved from the plan's buffer.  The layout depends only on
// (plan, nslices), so every call of one iteration sees the same addresses.
struct Scratch {
    float* xT;     // nslices transposed images (distance-driven forward)
    double* pf;    // forward-kernel FP64 partials [slice][block][3]
    double* pb;    // back-kernel FP64 partials    [slice][block][3]
    float* img;    // nslices images (ray-driven scatter target)
};
int get_scratch(const Plan& p, int nslices, Scratch* s, hipStream_t st);
// Bytes of scratch get_scratch carves for (plan, nslices) with the plan's current shapes.
size_t scratch_total(const Plan& p, int nslices);

// A caller-owned device region that get_scratch uses instead of the plan's buffer while
// it is set on the calling thread (graph capture gives each graph its own region).  A
// call whose layout does not fit fails with ITOMO_EINVAL and allocates nothing.
struct ScratchRegion {
    void* base;
    size_t bytes;
};
// Installs `r` on this thread for the lifetime of the guard.
class ScratchScope {
public:
    explicit ScratchScope(ScratchRegion r);
    ~ScratchScope();
    ScratchScope(const ScratchScope&) = delete;
    ScratchScope& operator=(const ScratchScope&) = delete;
private:
    ScratchRegion prev_;
};

// error reporting
int set_error(int status, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

// launchers (dd.hip / rd.hip / iter.hip)
// Tabulates Plan::d_lstar (allocates it; synchronous).
int dd_build_line_ends(Plan& p);
// Splits the views by forward launch into Plan::d_vlist (allocates it; synchronous).
int dd_build_view_lists(Plan& p);
int dd_forward(const Plan& p, const float* image, float* sino, int nslices,
               int64_t img_stride, int64_t sino_stride, int accumulate, hipStream_t st);
int dd_back(const Plan& p, const float* sino, float* image, int nslices,
            int64_t sino_stride, int64_t img_stride, int accumulate, hipStream_t st);
int rd_forward(const Plan& p, const float* image, float* sino, int nslices,
               int64_t img_stride, int64_t sino_stride, int accumulate, hipStream_t st);
int rd_back(const Plan& p, const float* sino, float* image, int nslices,
            int64_t sino_stride, int64_t img_stride, int accumulate, hipStream_t st);
int pd_forward(const Plan& p, const float* image, float* sino, int nslices,
               int64_t img_stride, int64_t sino_stride, int accumulate, hipStream_t st);
int pd_back(const Plan& p, const float* sino, float* image, int nslices,
            int64_t sino_stride, int64_t img_stride, int accumulate, hipStream_t st);
int ss_forward(const Plan& p, const float* image, float* sino, int nslices,
               int64_t img_stride, int64_t sino_stride, int accumulate, hipStream_t st);
int ss_back(const Plan& p, const float* sino, float* image, int nslices,
            int64_t sino_stride, int64_t img_stride, int accumulate, hipStream_t st);

}  // namespace itomo

// The opaque handle of the C ABI is the plan itself.
struct itomo_plan : itomo::Plan {};
