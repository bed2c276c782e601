// host.hpp -- host-side objects of the HIP deconvolution library.
#pragma once
#include <string>
#include <vector>

#include "engine.hpp"

namespace mvr {

// thread-local last error (read through mvr_last_error)
void set_error(const std::string& msg);
const char* last_error();

#define MVR_HIP(call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                                   \
      return -1;                                                                                      \
    }                                                                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// One FFT-convolution plan: output region D, kernel size K (see engine.hpp).
// ---------------------------------------------------------------------------
struct Plan {
  Dom dom{};
  int K[3] = {0, 0, 0};
  int M = 0, Hx = 0;
  float2* twM = nullptr;  // length-M twiddles (x)
  float2* wr = nullptr;   // exp(-2 pi i k / P0), k = 0..M
  float2* twY = nullptr;
  float2* twZ = nullptr;
  XFwdFn xf = nullptr;
  XRatioFn xr = nullptr;
  XUpdateFn xu = nullptr;
  XStoreFn xs = nullptr;
  StridedFn sy = nullptr, sz = nullptr;
  int wy = 0, wz = 0;  // tile widths of the y and z passes (0 = default)

  // chooses the FFT lengths; P[d] >= D[d] + K[d] - 1 (or the given minimum per axis)
  int create(const int D[3], const int K[3], const int* Pmin = nullptr);
  // the same geometry and kernel choice without device memory (twiddle pointers stay null): for callers that
  // place the twiddle tables themselves
  int create_geometry(const int D[3], const int K[3], const int* Pmin = nullptr);
  void destroy();
  size_t ncpx() const { return (size_t)dom.HxP * dom.P[1] * dom.P[2]; }
  size_t rows_inner() const { return (size_t)dom.D[1] * dom.D[2]; }
  StridedGeom geom_y() const;
  StridedGeom geom_z() const;
  // Narrows the geometries gy, gz of a convolution's y_conv and z_inv passes to what the next x pass reads: the
  // grid rows (y, z) in [m - lo, m + D + hi) per axis (within the grid).  y_conv then stores only those y, and
  // z_inv transforms only those y planes and stores only those z.  Batched geometries stay batched.
  void restrict_rows(const int lo[3], const int hi[3], StridedGeom& gy, StridedGeom& gz) const;
  // Sets q.skipOnes, the quotient rows q.rowLo / rowN, and the z geometry gz of the forward pass that reads the
  // quotient pass's result: the rows outside the quotient region are rows of ones (not written by the quotient
  // pass, put in by the z pass).  With gy (the y convolution after that z pass) the z pass also leaves out the
  // y planes outside the quotient region, planes of ones whose spectrum, P0 P2 at x and z frequency 0, the y
  // convolution puts in.  Not for batched blocks, whose quotient regions differ per block.
  void skip_ones_rows(RatioParams& q, StridedGeom& gz, StridedGeom* gy = nullptr) const;
  // For a forward transform of a mirror gather (g.mode MIRROR): sets g.uniqueRows, so that x_fwd transforms only
  // the rows that are not mirror copies, and makes the z pass after it (gzf) transform only those y planes and
  // read z through the mirror, and the y convolution after that (gyc) read y through it.  Not for batched
  // blocks, whose gathers mirror the whole volume, not the block; a no-op for spectra of 4 GiB or more.
  void mirror_rows(GatherParams& g, StridedGeom& gzf, StridedGeom& gyc) const;
  float inv_scale() const { return 2.0f / ((float)dom.P[0] * (float)dom.P[1] * (float)dom.P[2]); }

  // spectrum of a device kernel (K layout), scaled by the inverse-FFT factor
  int kernel_spectrum(const float* dKernel, float2* spec, hipStream_t s) const;
  // out (compact D) = K (*) ext(src) with src a volume of dims N; org = volume coord of FFT position 0
  int convolve(const GatherParams& g, const float2* spec, float2* work, float* out, hipStream_t s) const;
};

// DeconViewPSF.init for all views in list order (D/DeconViewPSF.java:119-254,
// D/DeconViews.java:69-70).  k1 are normalised in place; k2 receives the
// second kernel per view.  Compound kernels are convolved on the GPU.
int init_psfs(int dev, std::vector<std::vector<float>>& k1, const int K[3], int psfType, int nT,
              std::vector<std::vector<float>>& k2, hipStream_t s);
double sum_img_quirk(const float* a, size_t n, int nT);
void norm_to_sum1(float* a, size_t n, int nT);
void mirror_all(float* a, const int dim[3]);

enum { PSF_OPT_II = 0, PSF_OPT_I = 1, PSF_EFFICIENT_BAYESIAN = 2, PSF_INDEPENDENT = 3 };
enum { SCHEME_SEQ = 0, SCHEME_MUL = 1 };
enum { PSI_INIT_AVG_PRECISE = 0, PSI_INIT_FROM_ARRAY = 1, PSI_INIT_AVG_APPROX = 2, PSI_INIT_BLURRED_FUSED = 3 };

// per-pass timing categories (resident driver)
enum Pass { PASS_X_FWD = 0, PASS_Z_FWD, PASS_Y_CONV, PASS_Z_INV, PASS_X_RATIO, PASS_X_UPDATE, PASS_X_STORE,
            PASS_UPDATE_MUL, PASS_REDUCE, PASS_COUNT };

struct Timer {
  bool on = false, pending = false;
  unsigned mask = 0;
  int every = 1;                      // time one launch in `every` of each enabled category
  long long seen[PASS_COUNT] = {0};   // launches of each category since the last reset
  std::vector<hipEvent_t> ev;
  std::vector<int> cat;
  size_t used = 0;
  void begin(hipStream_t s, int c);
  void end(hipStream_t s);
  // a start / stop pair for a launch that records its own kernel's execution (both null when c is off)
  void kernel(int c, hipEvent_t* start, hipEvent_t* stop);
  int collect(double* ms, long long* n);  // accumulates into arrays of PASS_COUNT
  void release();
  // enable the categories in mask (bit c = category c), timing the first launch of every group of `every`;
  // clears the totals
  void reset(unsigned mask, int every = 1);
  double ms[PASS_COUNT] = {0};
  long long cnt[PASS_COUNT] = {0};
};

// ---------------------------------------------------------------------------
// Resident multiview deconvolution (MultiViewDeconvolution{Seq,Mul} over one
// domain covering the whole psi volume; D/MultiViewDeconvolution.java:90-212).
// ---------------------------------------------------------------------------
struct Decon {
  int dev = 0;
  hipStream_t stream = nullptr;
  int N[3] = {0, 0, 0};
  int V = 0, psfType = 0, scheme = 0, nT = 8;
  float lambda = 0.f, minValue = 1e-4f;
  int K[3] = {0, 0, 0};
  std::vector<float*> img, w;
  std::vector<std::vector<float>> kraw, k1, k2;
  std::vector<std::vector<float>> kfin1, kfin2;  // final kernels handed over by the caller (empty: none)
  float* psi = nullptr;
  float* psi0 = nullptr;
  Plan plan;
  std::vector<float2*> spec1, spec2;
  float2* S = nullptr;
  float2* Spsi = nullptr;
  float* T2 = nullptr;
  double2* partial = nullptr;
  int partialCap = 0;
  double2* statsDev = nullptr;
  int statsCap = 0;
  std::vector<double> stats;
  std::vector<float> maxI;
  double avg = 0;
  unsigned long long psiCount = 0;  // voxels with a positive view value (AvgPrecise pass)
  bool ready = false;
  Timer timer;

  // distance to a reference volume (convergence checks)
  float* ref = nullptr;
  double2* relDev = nullptr;

  // one iteration captured as a HIP graph and replayed (kernel arguments, including max[] and the
  // spectra pointers, are baked in: init and autotune drop the graph)
  bool useGraph = false;
  hipGraphExec_t graphExec = nullptr;
  double2* graphStats = nullptr;  // per-view (Seq) or per-iteration (Mul) stats of the replayed iteration
  // The reference's block structure (DeconView's block size B, D/DeconView.java:118-184: effective size
  // E = B - 2K + 2, margin K - 1; BlockGeneratorFixedSizePrecise), all blocks of one view in one launch per
  // pass.  blockB[0] == 0: one frame over the whole volume.
  int blockB[3] = {0, 0, 0};
  int blockE[3] = {0, 0, 0};
  int nbAll = 0;
  std::vector<int*> btab;   // per view: device table {o[3], D[3]} of its blocks with content
  std::vector<int> nbView;  // blocks per view after DeconView.filterBlocksForContent
  bool blocked() const { return blockB[0] != 0; }
  int set_block_size(const int B[3]);
  int build_blocks();
  void free_blocks();
  size_t partials_needed() const;
  // FFT lengths chosen by autotune (0 = the smallest compiled length)
  int tunedP[3] = {0, 0, 0};
  int tunedW[2] = {0, 0};  // tile widths of the y and z passes chosen by autotune (0 = default)

  int create(int dev, const int N[3], int V, int psfType, int scheme, float lambda, float minValue, int nT);
  int set_view(int v, const float* img, const float* w, const float* kernel, const int K[3]);
  int set_view_kernels(int v, const float* k1, const float* k2, const int K[3]);
  int set_max(const float* mx);
  int init(int psiInit, const float* psiIn);
  // psi0 := the psi currently resident (after an external initialiser wrote it)
  int commit_psi0();
  int run(int iterations);
  int run_view(int v);
  int reset_psi();
  int set_reference(const float* r);
  int rel_l2(double* out);
  int run_until(int maxIterations, double tol, int* iterationsOut, double* relOut);
  // FFT length autotuning: per axis, the `candidates` smallest compiled lengths >= D + K - 1 are timed on
  // view updates (coordinate descent x, y, z); the fastest set is kept and used by later init calls
  int autotune(int candidates, int chosen[3], double* msPerViewUpdate);
  int set_graph(bool on);
  void destroy();

  size_t nvox() const { return (size_t)N[0] * N[1] * N[2]; }
  int ensure_partials(int n);
  int build_plan_and_spectra(const int* Pmin, int nviews);
  void drop_graph();
  // one Seq view update; slot < 0: its statistics reduced into statOut now; slot >= 0: its partials left at
  // partial + slot * partials_needed() (count in slotCount[slot]) for reduce_slots
  int view_seq(int v, double2* statOut, int slot = -1);
  int slotCount[32] = {0};
  void reduce_slots(int n, double2* out);
  int iteration_mul(double2* statOut);
};

// ---------------------------------------------------------------------------
// Per-block compute thread (ComputeBlockSeqThread / ComputeBlockMulThread SPI,
// D/iteration/sequential/ComputeBlockSeqThreadCPU.java:53-169,
// D/iteration/mul/ComputeBlockMulThreadCPU.java:87-188).
// ---------------------------------------------------------------------------
struct BlockThread {
  int dev = 0, id = 0;
  hipStream_t stream = nullptr;
  int B[3] = {0, 0, 0};
  float lambda = 0.f, minValue = 1e-4f;
  int K[3] = {0, 0, 0};
  Plan plan;
  std::vector<float2*> spec1, spec2;
  float *psiB = nullptr, *imgB = nullptr, *wB = nullptr, *T2 = nullptr;
  std::vector<float*> imgV, wV;
  float2* S = nullptr;
  float2* Spsi = nullptr;
  double2* partial = nullptr;
  int partialCap = 0;
  double2* statDev = nullptr;

  int create(int dev, const int B[3], float minValue, float lambda, int id);
  int set_view_kernels(int v, const float* K1, const float* K2, const int K[3]);
  // whole view image and weights resident on the device (dims VN, shared by all views)
  std::vector<float*> viewImg, viewW;
  int VN[3] = {0, 0, 0};

  int set_view_data(int v, const float* img, const float* w, const int n[3]);
  int run_seq(int v, float* psiBlock, const float* imgBlock, const float* wBlock, float maxI, double stats[2]);
  int run_seq_at(int v, float* psiBlock, const long long off[3], float maxI, double stats[2]);
  int run_mul(int nv, float* psiBlock, const float* const* imgBlocks, const float* const* wBlocks,
              const float* maxIs, double stats[2]);
  int run_mul_at(int nv, float* psiBlock, const long long off[3], const float* maxIs, double stats[2]);
  void destroy();
  int ensure_partials(int n);
  int iterate_seq(int v, float maxI);
  int iterate_mul(int nv, const float* maxIs);
  int prepare_mul(int nv);
  int finish(float* psiBlock, double stats[2]);
};

}  // namespace mvr
