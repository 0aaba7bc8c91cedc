// Host-side launch interface of the device kernels.
#pragma once

#include <vector>

#include "dwt_common.h"

namespace dwt {

// ---- generic line kernels (line_kernels.hip) --------------------------------
template <typename T>
void launch_line_pass(bool fwd, const char* a, const char* b, char* c, long long line_step,
                      long long es, int nlines, int N, const Coeffs<T>& cf, T* gtmp,
                      int gtmp_blocks, hipStream_t st);
template <typename T>
void launch_zero_pad(char* ptr, long long sx, long long sy, int fw, int fh, const int zx[4],
                     const int zy[4], hipStream_t st);
int line_lds_limit_bytes();
hipError_t configure_line_kernels();

// ---- level-0 detail copy carried by other launches ----------------------------
// The in-place drivers move level 0's detail subbands between the image and
// scratch (forward: scratch -> image after level 0; inverse: image -> scratch
// before it).  Instead of a launch of its own, the copy rides on the launches
// of the small levels, which leave most of the GPU idle: a launch carrying
// rows [r0, r1) of the copy gets `blocks` extra workgroups that move them.
// The rows form one list: row y < ha is columns [xa, xa + wa) of image row y,
// row ha <= y < ha + hb is columns [0, wb) of image row y.  vec: 16-byte
// access allowed (both pointers, pitches and xa suitably aligned).
template <typename T>
struct RiderCopy {
  const T* src;
  long long sp;
  T* dst;
  long long dp;
  int xa, wa, ha, wb, hb;
  int r0, r1;
  int blocks;
  int vec;
  const int* rows;  // device list of image rows (entry i is row rows[i]); null: row i
};

// ---- in-place level 0 (inplace.hip) -------------------------------------------
// Plan of one level-0 launch on a W x H image (tiles of S-row segments, four
// strips per workgroup): the segment order, per segment the parts that fall
// back to scratch (forward: bit 0 HL rows, bit 1 LH + HH rows; inverse: bit 0
// the band is preserved in D), the tiles of each XCD in order (list[off[x] ..
// off[x+1]) for workgroups b = 8 k + x), every tile's dependencies (tiles
// whose reads must be done before it stores; CSR), and the image rows the
// riders copy (forward: D -> image after level 0; inverse: image -> D
// before it).  grid: workgroups of the launch.
struct L0Plan {
  bool fwd = true;
  int W = 0, H = 0, nsteps = 4, S = 8, band = 8;
  int nstrips = 0, ngroups = 0, nseg = 0, grid = 0;
  std::vector<int> order, segflags, list, list_off, dep_off, deps, rows;
};
L0Plan build_l0_plan(bool fwd, int W, int H, int nsteps, int S);
extern int g_l0_band;
// Device view of a plan: tables in device memory, the flags (one word per
// tile) and error word of the scratch set, this launch's epoch.
struct L0Dev {
  const int* list;
  const int* dep_off;
  const int* deps;
  const int* segflags;
  unsigned* flags;
  unsigned* err;
  unsigned long long spin_limit;
  unsigned epoch;
  int off[9];
  int ngroups, nstrips;
};
// Forward level 0 of the W x H image A (pitch ap) in place: LL to ll, details
// to A (or D, pitch dp, for the parts that fall back).  Inverse: LL from ll,
// details from A (D for preserved bands), output to A.
template <typename T>
void launch_fwd_l0(T* A, long long ap, int W, int H, T* ll, long long llp, T* D, long long dp, const Coeffs<T>& cf,
                   const L0Dev& ip, int S, int grid, hipStream_t st);
template <typename T>
void launch_inv_l0(const T* ll, long long llp, T* A, long long ap, const T* D, long long dp, int W, int H,
                   const Coeffs<T>& cf, const L0Dev& ip, int S, int grid, hipStream_t st);
// A rider copy as a launch of its own (rc.blocks workgroups).
template <typename T>
void launch_rider_alone(const RiderCopy<T>& rc, hipStream_t st);

// ---- packed fast path (tile_kernels.hip) ------------------------------------
// Forward level on a packed W x H region `src` (row pitch sp elements):
// LL (ceil(W/2) x ceil(H/2)) goes to `ll` (pitch llp); HL, LH, HH go to `det`
// (pitch dp) at their Mallat positions.  src must not overlap ll or det.
template <typename T>
void launch_fwd_level(const T* src, long long sp, int W, int H, T* ll, long long llp, T* det,
                      long long dp, const Coeffs<T>& cf, hipStream_t st,
                      const RiderCopy<T>* rider = nullptr);
// Inverse level: LL from `ll` (pitch llp), HL/LH/HH from `det` (pitch dp) at
// their Mallat positions of the W x H fine region; output to `dst` (pitch
// dstp).  dst must not overlap ll or the details read from det.
template <typename T>
void launch_inv_level(const T* ll, long long llp, const T* det, long long dp, T* dst,
                      long long dstp, int W, int H, const Coeffs<T>& cf, hipStream_t st,
                      const RiderCopy<T>* rider = nullptr);
// Tile-shape table of the level kernels, by direction (1 forward, 0 inverse)
// and height class ceil(log2(H)): S rows per segment, U prefetched row pairs
// (S = 0 restores the built-in heuristic).  tile_get reports the shape a
// launch of that size would use.
int tile_set(int fwd, int hclass, int S, int U);
int tile_get(int fwd, int H, int W, int nsteps, int* S, int* U);
// Block kernels: a whole small pyramid in one workgroup's LDS.  Forward:
// levels 0..nlev-1 of the W0 x H0 region read from src, written in Mallat
// layout to dst.  Inverse: the nlev levels whose finest output is W0 x H0,
// coefficients from src, output to dst.  src may equal dst (the region is read
// completely before it is written).  tail_start picks the first level of an
// L-level transform that the block kernel takes (L when none).
template <typename T>
void launch_block_fwd(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st, const RiderCopy<T>* rider = nullptr);
template <typename T>
void launch_block_inv(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st, const RiderCopy<T>* rider = nullptr);
int tail_start(int W, int H, int L, int elem);
// Sparse frames (outer Wo x Ho, inner Wi x Hi, contiguous rows): levels
// [sparse_tail_start, L) as one block launch, in place on the region at the
// frame's origin (outer W0 x H0 and inner Wi0 x Hi0 at that level).
int sparse_tail_start(int Wo, int Ho, int Wi, int Hi, int L, int elem);
// One level of a sparse frame (outer Wo x Ho, inner Wi x Hi, both inner sides
// >= 2) in one launch (sparse.hip).  Forward: in -> q (the level's whole
// output region); rider: a copy carried by extra workgroups.  Inverse: the
// previous level's output prev (covering [0, ceil(Ho/2)) x [0, ceil(Wo/2))) and
// the frame's details (everything else) -> q; rider as in the forward.
template <typename T>
void launch_sparse_fwd(const T* in, long long ip, T* q, long long qp, int Wo, int Ho, int Wi, int Hi, int zp,
                       const Coeffs<T>& cf, const RiderCopy<T>* rider, hipStream_t st);
template <typename T>
void launch_sparse_inv(const T* prev, long long pp, const T* frame, long long fp, T* q, long long qp, int Wo,
                       int Ho, int Wi, int Hi, int zp, const Coeffs<T>& cf, const RiderCopy<T>* rider,
                       hipStream_t st);
template <typename T>
void launch_block_sparse(bool fwd, T* a, long long ap, int W0, int H0, int Wi0, int Hi0, int nlev, int zp,
                         const Coeffs<T>& cf, hipStream_t st);
void tail_config(long long max_region_bytes, int unused);
hipError_t configure_tail_kernels();
// Fused middle levels (mid_kernels.hip): nl (<= kMidMaxLevels) consecutive
// levels in one launch, each workgroup computing overlapping LDS windows.
// Forward: LL_j (W0 x H0) from src; details of the nl levels into det (Mallat
// layout, in place); LL_{j+nl} to ll.  Inverse: LL_{j+nl} from llsrc,
// details from det; LL_j (W0 x H0) to out.
constexpr int kMidMaxLevels = 4;
template <typename T>
void launch_mid_fwd(const T* src, long long sp, T* det, long long dp, T* ll, long long llp, int W0,
                    int H0, int nl, const Coeffs<T>& cf, hipStream_t st);
template <typename T>
void launch_mid_inv(const T* llsrc, long long lp, const T* det, long long dp, T* out, long long op,
                    int W0, int H0, int nl, const Coeffs<T>& cf, hipStream_t st);
size_t mid_lds_bytes(bool fwd, int nl, int nsteps, int elem);
hipError_t configure_mid_kernels();
// Raise the batched tail kernels' dynamic LDS limit (batch.hip).
hipError_t configure_batch_tail();
// Device-side compare of two images (verify.hip): scratch of
// compare_scratch_bytes(W * H) bytes; the result, compare_result_bytes() at
// compare_result_offset(W * H), is decoded by compare_result after a copy to
// the host.  eps: the threshold of the differ verdict.
constexpr int kUlpBins = 32;
struct CompareResult {
  double max_abs, sum_sq, peak_b;
  long long count, nonfinite, over;
  long long ulp[kUlpBins];
};
long long compare_blocks(long long n);
size_t compare_scratch_bytes(long long n);
size_t compare_result_bytes();
size_t compare_result_offset(long long n);
void launch_compare(bool a64, const void* a, long long sxa, long long sya, bool b64, const void* b,
                    long long sxb, long long syb, int W, int H, double eps, void* scratch,
                    hipStream_t st);
void compare_result(const void* host_copy, CompareResult* r);
// FP32 image (byte strides) to a packed FP64 one.
void launch_to_f64(const void* src, long long sx, long long sy, double* dst, int W, int H,
                   hipStream_t st);
// Copy a W x H frame between element-strided layouts (row stride r, column
// stride c, in elements).
template <typename T>
void launch_copy_strided(const T* src, long long sr, long long sc, T* dst, long long dr,
                         long long dc, int W, int H, hipStream_t st);
// The C channels of a W x H image (element strides: rows sr, pixels sx,
// channels sc) to (gather) or from C packed planes, plane c at
// planes + c * plane_step with row pitch pp; every pixel's channels in one
// pass.
template <typename T>
void launch_channels(bool gather, T* img, long long sr, long long sx, long long sc, int C, int W, int H,
                     T* planes, long long pp, long long plane_step, hipStream_t st);
// Copy a w x h rectangle between packed buffers.
template <typename T>
void launch_copy_rect(const T* src, long long sp, T* dst, long long dp, int w, int h,
                      hipStream_t st);

// ---- pyramid kernels (pyr_kernels.hip) ----------------------------------------
// One level of one image in a pyramid launch.  Forward: src = level input
// (W x H), out = LL of the next level, det = detail subbands at their Mallat
// positions.  Inverse: src = LL of the coarser level, det = details at their
// Mallat positions of the W x H output region, out = that region.  Pitches in
// elements.  S = segment rows; nstrips .. cnt0 are filled by pyr_layout; dep
// is the index of the task that produces src in the same launch (-1: none;
// it must precede this task in the table).
struct PyrTask {
  const void* src;
  long long sp;
  void* out;
  long long op;
  void* det;
  long long dp;
  int W, H, S;
  int nstrips, ngroups, nseg;
  int item0, cnt0;
  int dep, reserved;
};
constexpr int kPyrMaxTasks = 32;  // tasks of one pyramid launch
// Device synchronisation state of a launch: ctrl[0] unused, ctrl[1]
// finished workgroups; cnt[ncnt] finished workgroups per (task, segment), all
// zero before a launch (the launch leaves them zero); deps[ncnt] producer
// segment range of each (task, segment); err set by a wait that ran longer
// than spin_limit ticks of the constant-rate clock.
// trace (optional, may be null): per item {start, dependencies met, done}
// ticks of the constant-rate clock and the workgroup index.
struct PyrSync {
  unsigned* ctrl;
  unsigned* cnt;
  const int2* deps;
  unsigned* err;
  int ncnt;
  unsigned long long spin_limit;
  unsigned long long* trace;
};
// Assign items and counters to tasks (in table order); returns the item count.
int pyr_layout(PyrTask* tasks, int ntasks, int nsteps);
// Producer segment range of every segment of task t (host).
void pyr_deps(const PyrTask& t, const PyrTask& producer, bool fwd, int nsteps, int2* out);
// Launch a pyramid over the device task table; all tasks use segment rows S
// (8, 16 or 32); vec: every src/out pointer and pitch allows two-element
// vector access.  grid_cap > 0 limits the number of workgroups.
template <typename T>
void launch_pyr(bool fwd, bool vec, int S, const PyrTask* d_tasks, int ntasks, int nitems,
                const PyrSync& sy, const Coeffs<T>& cf, int grid_cap, hipStream_t st);
void launch_fill_u32(unsigned* p, unsigned v, long long n, hipStream_t st);

// ---- full-depth 1D transforms of many signals (lines.hip) ---------------------------
// Signals as the rows of a matrix (count rows, contiguous samples): one level
// on n samples.  Forward: a = input rows (pitch ap), b = L out (pitch bp),
// c = H out (pitch cp, H sample k at c[k]).  Inverse: a = L in, b = H in
// (H sample k at b[k]), c = output.  vec: pair-wise access of the input
// (forward) or output (inverse) rows allowed.
template <typename T>
void launch_row_level(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                      long long cp, int n, int count, const Coeffs<T>& cf, bool vec,
                      hipStream_t st);
// The same level on whole rows held in one workgroup's LDS: every row is read
// before any output is written, so the outputs may overlap the input rows
// (in place).  Same argument meaning; returns false (nothing launched) when a
// row of n samples is too long to hold or shorter than 64 samples.
bool row_resident_fits(int n, int elem);
template <typename T>
bool launch_row_resident(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                         long long cp, int n, int count, const Coeffs<T>& cf, hipStream_t st);
// Signals as the columns of a matrix (count columns): one level on n rows.
// Forward: a = input rows (pitch ap), b = L rows out (pitch bp), c = H rows
// out from row ceil(n/2) (pitch cp).  Inverse: a = L rows, b = H rows from
// row ceil(n/2), c = output rows.  vec: pair-wise access allowed.
template <typename T>
void launch_col_level(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                      long long cp, int n, int count, const Coeffs<T>& cf, bool vec,
                      hipStream_t st);
// The same level restricted to segments [seg0, seg0 + nseg) of
// col_segment_rows() rows (col_segments(n) in all), with H row m at row
// hbase + m of the H buffer (launch_col_level: hbase = ceil(n/2)).  Forward
// segment s reads input rows [R s - NS, R (s + 1) + NS) and writes output
// pairs [R s / 2, R (s + 1) / 2); inverse segment s reads H rows
// [R s / 2 - NS / 2, R (s + 1) / 2 + NS / 2) and writes output rows
// [R s, R (s + 1)) (R = col_segment_rows(), NS = lifting steps).
int col_segments(int n);
int col_segment_rows();
void col_segment_access(bool fwd, int n, int nsteps, int seg, std::vector<int>& reads,
                        std::vector<int>& writes);
template <typename T>
void launch_col_segments(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                         long long cp, int n, int count, int seg0, int nseg, int hbase,
                         const Coeffs<T>& cf, bool vec, hipStream_t st);

}  // namespace dwt
