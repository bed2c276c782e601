/* mvr_hip.h -- C ABI of libmvr_hip.so, the MI355X (gfx950) multiview
 * Richardson-Lucy deconvolution path.
 *
 * Conventions (all entry points):
 *   - return 0 on success and a negative value on error; the message is in
 *     mvr_last_error() (thread-local);
 *   - arrays are float32 in ImgLib2 flat order (x fastest), dims are given
 *     as {x, y, z} unless the name says ZYX;
 *   - array arguments may be host memory or device memory of the handle's
 *     GPU (hipMemcpyDefault semantics); they are only borrowed for the call;
 *   - calls are synchronous; one handle must not be used by two threads at
 *     once, different handles may run concurrently (one per Java block thread).
 *
 * Citations are relative to src/main/java/net/preibisch/mvrecon/process/ of
 * the reference (PreibischLab/multiview-reconstruction).
 */
#ifndef MVR_HIP_H
#define MVR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* PSF types, enum order of DeconViewPSF.PSFTYPE (deconvolution/DeconViewPSF.java:52) */
#define MVR_PSF_OPTIMIZATION_II 0
#define MVR_PSF_OPTIMIZATION_I 1
#define MVR_PSF_EFFICIENT_BAYESIAN 2
#define MVR_PSF_INDEPENDENT 3
/* iteration schemes: MultiViewDeconvolutionSeq / MultiViewDeconvolutionMul */
#define MVR_SCHEME_SEQ 0
#define MVR_SCHEME_MUL 1
/* PsiInit: PsiInitAvgPrecise (deconvolution/init/PsiInitAvgPrecise.java:52-112),
 * PsiInitFromRAI with the precise max (deconvolution/init/PsiInitFromRAI.java:65-82) */
#define MVR_PSI_INIT_AVG_PRECISE 0
#define MVR_PSI_INIT_FROM_ARRAY 1
/* PsiInitAvgApprox (deconvolution/init/PsiInitAvgApprox.java:47-99, PsiInitAvgApproxThread.java:57-85):
 * avg and max[] from each view's middle x hyperslice */
#define MVR_PSI_INIT_AVG_APPROX 2
/* PsiInitBlurredFused (deconvolution/init/PsiInitBlurredFused.java:62-127,
 * deconvolution/util/FusedNonZeroRandomAccess.java:57-96): weighted fusion of the positive views blurred
 * with a Gaussian of sigma 5 (mirror-single borders); max[] and avg as AvgPrecise */
#define MVR_PSI_INIT_BLURRED_FUSED 3
/* extension modes of mvr_fft_convolve (util/FFTConvolution.java usage in
 * ComputeBlockSeqThreadCPU.java:171-209 and DeconViewPSF.java:152-225) */
#define MVR_EXT_MIRROR_SINGLE 0
#define MVR_EXT_VALUE 1

/* ---- device queries: replace CUDAStandardFunctions (cuda/CUDAStandardFunctions.java:27-46) ---- */
/* getNumDevicesCUDA: number of HIP devices, -1 if the runtime failed */
int mvr_get_num_devices(void);
/* getNameDeviceCUDA: device name into name[len] */
int mvr_get_device_name(int dev, char* name, int len);
/* getMemDeviceCUDA / getFreeMemDeviceCUDA: bytes, -1 on error */
long long mvr_get_mem_device(int dev);
long long mvr_get_free_mem_device(int dev);
/* getCUDAcomputeCapabilityMajor/MinorVersion: gfx architecture (9 / 50 for gfx950) */
int mvr_get_arch_major(int dev);
int mvr_get_arch_minor(int dev);

/* ---- replaces CUDAFourierConvolution.convolution3DfftCUDAInPlace (cuda/CUDAFourierConvolution.java:32),
 * called by ComputeBlockSeqThreadCUDA.convolve1/convolve2 (deconvolution/iteration/sequential/
 * ComputeBlockSeqThreadCUDA.java:171-208).  In-place 3D convolution of im with kernel, dims as {z,y,x}
 * like CUDATools.getCUDACoordinates.  The image is extended by mirror-single reflection (the CPU
 * path's convolve1); the closed CUDA library's border behaviour is not specified by the reference. */
int mvr_convolution3d_fft_inplace(float* im, const int* imDimZYX, const float* kernel, const int* kernelDimZYX,
                                  int dev);

/* ---- one util.FFTConvolution.convolve (util/FFTConvolution.java:490-506): out = kernel (*) ext(img) ---- */
int mvr_fft_convolve(int dev, const float* img, const int nXYZ[3], int ext, float extValue, const float* kernel,
                     const int kXYZ[3], float* out);
/* FFT lengths with compiled kernels, ascending, into out[0..cap); returns how many there are.  A domain of
 * extent n + k - 1 along an axis runs at the smallest of them >= that extent (even lengths only along x).
 * The reference picks its padded sizes inside imglib2's FFTMethods; this list has no counterpart there. */
int mvr_fft_lengths(int* out, int cap);

/* ---- mvr_fft_convolve without allocation.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  img, kernel and out must be device memory
 * of `dev` (refused otherwise); the twiddle tables, the kernel spectrum and the work spectrum live in one static
 * device array of MVR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, so FFT grids up to about 180^3
 * fit (e.g. a 128^3 block with a 41^3 kernel); larger ones are refused with an error before any device work.
 * The kernels are enqueued on `stream` (a hipStream_t; NULL = the default stream) and the call returns without
 * waiting: synchronise the stream before reading out.  All calls on one GPU share its array, so calls that
 * overlap on one GPU race: serialise them (one stream). ---- */
#define MVR_NOALLOC_SCRATCH_BYTES (64ll << 20)
int mvr_fft_convolve_noalloc(int dev, const float* img, const int nXYZ[3], int ext, float extValue,
                             const float* kernel, const int kXYZ[3], float* out, void* stream);

/* ---- DeconvolutionMethods on whole arrays (deconvolution/iteration/sequential/DeconvolutionMethods.java): the
 * elementwise steps that the deconvolution passes fuse into their transforms, run alone over n values in flat order
 * (same float and double rounding as the Java methods, no FMA contraction).  stats = {sumChange, maxChange} of
 * IterationStatistics; sumChange is summed in double in the library's order (the reference sums per thread portion). */
/* computeQuotient (:46-98): psiBlurred[i] = img[i] > 0 ? img[i] / psiBlurred[i] : 1 */
int mvr_compute_quotient(int dev, const float* img, float* psiBlurred, long long n);
/* computeFinalValues (:103-182) with computeNextValue (:320-358) and tikhonov (:421): psi updated in place */
int mvr_compute_final_values(int dev, float* psi, const float* integral, const float* weight, long long n,
                             double lambda, float minIntensity, float maxIntensity, double stats[2]);
/* computeFinalValuesMul (:187-305) with computeNextValueMul (:370-419): numViews integrals and weights, n < 2^31 */
int mvr_compute_final_values_mul(int dev, float* psi, int numViews, const float* const* integral,
                                 const float* const* weight, long long n, double lambda, float minIntensity,
                                 float maxIntensity, double stats[2]);

/* ---- DeconViewPSF.init for all views in list order (deconvolution/DeconViewPSF.java:119-254,
 * deconvolution/DeconViews.java:69-70): kernel1 normalised to sum 1 (AdjustInput.normToSum1 with the
 * numeric thread count nT), kernel2 per PSF type; compound-kernel convolutions run on the GPU.
 * kernels/k1Out/k2Out hold numViews kernels of kXYZ each, back to back. ---- */
int mvr_init_psfs(int dev, const float* kernels, int numViews, const int kXYZ[3], int psfType, int nT, float* k1Out,
                  float* k2Out);

/* ---- per-block compute thread: replaces ComputeBlockSeqThreadCPU / ComputeBlockMulThreadCPU
 * (deconvolution/iteration/sequential/ComputeBlockSeqThreadCPU.java:53-169,
 *  deconvolution/iteration/mul/ComputeBlockMulThreadCPU.java:87-188), one per Java block thread. ---- */
typedef struct mvr_block_thread mvr_block_thread;
/* constructor (minValue, lambda, id, blockSize) */
int mvr_block_thread_create(mvr_block_thread** out, int dev, const int blockSizeXYZ[3], float minValue, float lambda,
                            int id);
/* DeconViewPSF.getKernel1FFT/getKernel2FFT (deconvolution/DeconViewPSF.java:79-111): spectra of the
 * view's final kernel1/kernel2 at block size, cached per view id */
int mvr_block_thread_set_view_kernels(mvr_block_thread* t, int view, const float* kernel1, const float* kernel2,
                                      const int kXYZ[3]);
/* ComputeBlockSeqThread.runIteration (iteration/sequential/ComputeBlockSeqThread.java:54-61): psiBlock
 * (= getPsiBlockTmp, mirror-extended psi) is updated in place; imgBlock/weightBlock are the zero-extended
 * view image and weights over the block; stats = {sumChange, maxChange} (IterationStatistics) */
int mvr_block_run_iteration_seq(mvr_block_thread* t, int view, float* psiBlock, const float* imgBlock,
                                const float* weightBlock, float maxIntensityView, double stats[2]);
/* ComputeBlockMulThread.runIteration (iteration/mul/ComputeBlockMulThread.java:56-62) over views 0..numViews-1 */
int mvr_block_run_iteration_mul(mvr_block_thread* t, int numViews, float* psiBlock, const float* const* imgBlocks,
                                const float* const* weightBlocks, const float* maxIntensityViews, double stats[2]);
/* The view's whole image and weights (nXYZ, the same for every view), kept on the device for the thread's
 * lifetime.  The reference's drivers cut a zero-extended image and weight block from them for every block in every
 * iteration (MultiViewDeconvolutionSeq.java:115-126, MultiViewDeconvolutionMul.java:173-183); with the data
 * resident, the _at entry points cut those blocks on the device and only psiBlock crosses the bus. */
int mvr_block_thread_set_view_data(mvr_block_thread* t, int view, const float* img, const float* weight,
                                   const int nXYZ[3]);
/* mvr_block_run_iteration_seq with imgBlock / weightBlock = the zero-extended view data over the block whose first
 * voxel is at volume coordinate offsetXYZ (the Block interval min, cuda/Block.java:91-133; may be negative) */
int mvr_block_run_iteration_seq_at(mvr_block_thread* t, int view, float* psiBlock, const long long offsetXYZ[3],
                                   float maxIntensityView, double stats[2]);
/* mvr_block_run_iteration_mul with the blocks of views 0..numViews-1 cut from their resident data */
int mvr_block_run_iteration_mul_at(mvr_block_thread* t, int numViews, float* psiBlock, const long long offsetXYZ[3],
                                   const float* maxIntensityViews, double stats[2]);
void mvr_block_thread_destroy(mvr_block_thread* t);

/* ---- resident deconvolution: the whole MultiViewDeconvolution{Seq,Mul} run on one GPU
 * (deconvolution/MultiViewDeconvolution.java:90-212, MultiViewDeconvolutionSeq.java:58-180,
 *  MultiViewDeconvolutionMul.java:116-245).  psi, views, weights and kernel spectra stay in HBM; the
 * psi volume is treated as one block whose faces are mirror-extended, which equals the reference's
 * blocked result on every pasted voxel (margin K-1 per block). ---- */
typedef struct mvr_decon mvr_decon;
/* DeconViews + MultiViewDeconvolution ctor parameters; nT is the numeric thread count of the
 * kernel normalisation (Threads.numThreads(), Threads.java:40) */
int mvr_decon_create(mvr_decon** out, int dev, const int psiDimXYZ[3], int numViews, int psfType, int scheme,
                     float lambda, float minValue, int nT);
/* DeconView(image, weight, kernel) (deconvolution/DeconView.java:118-184) */
int mvr_decon_set_view(mvr_decon* d, int view, const float* img, const float* weight, const float* kernel,
                       const int kXYZ[3]);
/* DeconViewPSF.getKernel1 / getKernel2 (deconvolution/DeconViewPSF.java:76-77) of a view whose DeconViewPSF.init
 * already ran on the caller's side (the DeconViews ctor does it, DeconViews.java:69-70; kernel1 is then
 * normalised in place): when every view has them, mvr_decon_init uses them as they are instead of running
 * DeconViewPSF.init on the raw kernels of mvr_decon_set_view. */
int mvr_decon_set_view_kernels(mvr_decon* d, int view, const float* kernel1, const float* kernel2,
                               const int kXYZ[3]);
/* The reference's block structure instead of one frame: DeconView's block size (deconvolution/DeconView.java:
 * 118-184; effective size B - 2K + 2 per axis, margin K - 1, cuda/BlockGeneratorFixedSizePrecise.java:59-131, uneven
 * tail blocks), blocks without weights dropped per view (DeconView.filterBlocksForContent, :204-274).  All blocks
 * of one view run in one launch per pass (they share the FFT grid and the kernel spectra).  Seq scheme; the
 * effective size must be at least K - 1 (the Seq driver's write-back lag then makes every view pass act on the
 * psi before it, MultiViewDeconvolutionSeq.java:150-157, as the batched launch does).  {0, 0, 0}: one frame
 * (default).  Takes effect at the next mvr_decon_init. */
int mvr_decon_set_block_size(mvr_decon* d, const int blockSizeXYZ[3]);
/* blocks in the volume (returned) and per view the blocks with content (blocksPerView[0..numViews)) */
int mvr_decon_get_blocks(mvr_decon* d, int* blocksPerView, int cap);
/* DeconViewPSF.init for all views + PsiInit.runInitialization (MultiViewDeconvolution.java:102-135) */
int mvr_decon_init(mvr_decon* d, int psiInit, const float* psi0);
/* PsiInit.getMax of an initialiser that ran on the caller's side (MultiViewDeconvolution.java:124-134), after
 * mvr_decon_init(MVR_PSI_INIT_FROM_ARRAY, psi): the per-view maxima that normalise the Tikhonov term */
int mvr_decon_set_max(mvr_decon* d, const float* maxPerView);
/* runIterations: `iterations` more iterations */
int mvr_decon_run(mvr_decon* d, int iterations);
/* one view pass of MultiViewDeconvolutionSeq.runNextIteration (MultiViewDeconvolutionSeq.java:69-162): the update
 * of psi by view `view` alone (Seq scheme only); its statistics are appended like a run's */
int mvr_decon_run_view(mvr_decon* d, int view);
int mvr_decon_get_psi(mvr_decon* d, float* out);
/* PsiInit.getMax / getAvg */
int mvr_decon_get_max(mvr_decon* d, float* maxOut, double* avgOut);
/* DeconViewPSF.getKernel1/getKernel2 after init */
int mvr_decon_get_kernels(mvr_decon* d, int view, float* kernel1, float* kernel2);
/* IterationStatistics {sumChange, maxChange} per view (Seq) or per iteration (Mul), in run order;
 * returns the number of doubles available */
int mvr_decon_get_stats(mvr_decon* d, double* out, int cap);
/* restore psi to the initial estimate (repeatable runs) */
int mvr_decon_reset_psi(mvr_decon* d);
/* FFT geometry: out[0..2] = FFT lengths, out[3] = half-spectrum row pitch */
int mvr_decon_get_geometry(mvr_decon* d, int out[4]);
/* per-pass GPU timing on the library's stream: the strided passes (z_fwd, y_conv, z_inv) with start and stop
 * events recorded by the launch itself (the kernel's own execution), the others with events recorded around
 * the launch.  mask bit c enables category c (0 x_fwd, 1 z_fwd, 2 y_conv, 3 z_inv, 4 x_ratio, 5 x_update,
 * 6 x_store, 7 update_mul, 8 reduce); 0 disables.  Totals are cleared by every call. */
int mvr_decon_profile(mvr_decon* d, unsigned mask);
/* the same, timing only the first launch of every group of `every` launches of each enabled category (every
 * >= 1): the events of a timed launch delay the stream by a few microseconds, so a throughput measured over
 * the same run is disturbed less; launches[c] counts the timed launches */
int mvr_decon_profile_sampled(mvr_decon* d, unsigned mask, int every);
int mvr_decon_pass_times(mvr_decon* d, double msTotal[9], long long launches[9]);
void mvr_decon_destroy(mvr_decon* d);

/* ---- input preparation (deconvolution/util/ProcessInputImages.java:279-399): each view resampled into the
 * deconvolution bounding box, its blending weights, and the weights normalised across views.
 * affine = the view's registration (input pixel -> world) as a 3 x 4 row-major matrix; the bounding box
 * is given by its world minimum bbMin and size bbXYZ; outputs are bbXYZ volumes. ---- */
/* TransformView.transformView (fusion/transformed/TransformView.java:59-73, TransformedInputRandomAccess.java:61-84):
 * outside the strict interior of the input -> outsideValue; inside, linear (interpolation 1) or nearest (0)
 * interpolation, clamped to >= minValue when hasMinValue (MultiViewDeconvolution.minValueImg = 1) */
int mvr_transform_view(int dev, const float* img, const int imgXYZ[3], const double affine[12], const long long bbMin[3],
                       const int bbXYZ[3], int hasMinValue, float minValue, float outsideValue, int interpolation,
                       float* out);
/* TransformWeight.transformBlending (fusion/transformed/TransformWeight.java:82-112, 114-134) with
 * BlendingRealRandomAccess.computeWeight (fusion/transformed/weights/BlendingRealRandomAccess.java:95-127)
 * on the input interval [0, imgXYZ - 1]; border and blending per dimension (after FusionTools.adjustBlending) */
int mvr_transform_blending(int dev, const int imgXYZ[3], const float border[3], const float blending[3],
                           const double affine[12], const long long bbMin[3], const int bbXYZ[3], float* out);
/* ProcessInputImages.fuseGroups for one group of numViews views (deconvolution/util/ProcessInputImages.java:
 * 279-399), in one pass without per-view intermediates: view j (imgXYZ[3j..], registration affines[12j..]) is
 * intensity-adjusted (fusion/intensityadjust/IntensityAdjuster.java:43-54: (float)(v * m00 + m01) with
 * intensityAdjust[2j..2j+1]; NULL = none) and resampled like mvr_transform_view (min 1, outside 0, linear);
 * imgOut = FusedRandomAccess AVG (fusion/transformed/FusedRandomAccess.java:67-88) weighted by the view's fusion
 * blending, weightOut = CombineWeightsSumRandomAccess (weightcombination/CombineWeightsSumRandomAccess.java:40-50)
 * of the views' deconvolution blending.  fusionBlending / deconBlending hold per view {border xyz, range xyz}
 * (after FusionTools.adjustBlending); NULL = the constant weight 1 of the reference's no-blending branch. */
int mvr_fuse_group(int dev, int numViews, const float* const* imgs, const int* imgXYZ, const double* affines,
                   const double* intensityAdjust, const float* fusionBlending, const float* deconBlending,
                   const long long bbMin[3], const int bbXYZ[3], float* imgOut, float* weightOut);
/* NormalizingRandomAccess.get for all views (deconvolution/normalization/NormalizingRandomAccess.java:76-108,
 * 183-214): weights[v] (nXYZ each) are replaced in place */
int mvr_normalize_weights(int dev, float* const* weights, int numViews, const int nXYZ[3], double osemSpeedup,
                          int additionalSmoothBlending, float maxDiffRange, float scalingRange);

/* ---- PSF preparation (deconvolution/util/PSFPreparation.java:41-89): PSFExtraction.getTransformedNormalizedPSF
 * (psf/PSFExtraction.java:182-193, 367-470): min-max normalisation, then resampling by the view's affine into an
 * odd-sized grid whose centre is the transformed PSF centre (linear interpolation, zero outside).
 * mvr_transform_psf_size gives the output dims; out holds that many floats. ---- */
int mvr_transform_psf_size(const int psfXYZ[3], const double affine[12], int outXYZ[3]);
int mvr_transform_psf(int dev, const float* psf, const int psfXYZ[3], const double affine[12], float* out);

/* ---- batching and scheduling of the resident run (no counterpart in the reference, whose iteration
 * loop is MultiViewDeconvolution.runIterations, deconvolution/MultiViewDeconvolution.java:144-200) ----
 * Autotuning: per axis the candidatesPerAxis smallest compiled FFT lengths >= extent + K - 1 are timed on
 * view updates (coordinate descent over x, y, z), then the tile width (8 or 16 columns) of the y and z passes;
 * a trial is the fastest of 3 timed view updates, and a candidate replaces the incumbent only when it is more
 * than 1 % faster.  The chosen schedule replaces the plan and is kept for later mvr_decon_init calls.  Must
 * follow mvr_decon_init; psi is restored to the initial estimate. */
int mvr_decon_autotune(mvr_decon* d, int candidatesPerAxis, int chosenXYZ[3], double* msPerUpdate);
/* enable != 0: each iteration (all views) is captured once as a HIP graph and replayed; the graph is
 * rebuilt after mvr_decon_init / mvr_decon_autotune.  Pass timing (mvr_decon_profile) runs eagerly. */
int mvr_decon_use_graph(mvr_decon* d, int enable);
/* tile widths (columns) of the y and z passes in use: 0 = the default for the length, else 8 or 16 as chosen
 * by mvr_decon_autotune or set here (a width not compiled for a length runs the default); the result does
 * not depend on the tiling */
int mvr_decon_get_tile_widths(mvr_decon* d, int out[2]);
int mvr_decon_set_tile_widths(mvr_decon* d, int wy, int wz);

/* ---- convergence checks against a reference volume (FP64 device reductions).  The reference keeps
 * no iteration-wise distance; its IterationStatistics (deconvolution/DeconvolutionMethods.java:308)
 * are returned by mvr_decon_get_stats.  The reference volume is typically a long plain RL run. */
int mvr_decon_set_reference(mvr_decon* d, const float* ref);
/* *out = ||psi - ref||_2 / ||ref||_2 */
int mvr_decon_rel_l2(mvr_decon* d, double* out);
/* run iterations until rel-L2 <= tol (checked after each iteration) or maxIterations are done */
int mvr_decon_run_until(mvr_decon* d, int maxIterations, double tol, int* iterationsOut, double* relOut);

/* ---- learned initial estimate: a new PsiInit (deconvolution/init/PsiInit.java:32-47) built like
 * PsiInitFromRAI (deconvolution/init/PsiInitFromRAI.java:65-82): psi comes from the network, max[]
 * and avg from the AvgPrecise pass.  The network (3-level 3D U-Net, 16/32/64 channels, 3^3
 * convolutions, ReLU; BF16 on MFMA with FP32 accumulation) and its training have no counterpart in
 * the reference.  Its inputs are the V weighted views and their mean; its output is psi. */
typedef struct mvr_unet mvr_unet;
/* views: 1..8 */
int mvr_unet_create(mvr_unet** out, int dev, int views, unsigned long long seed);
/* A trained network from its weights file (the .npz that mvr/learned.py writes: float32 vector `params`,
 * integer scalar `views`, stored uncompressed), so a caller only passes a path: the role of the weights path
 * of a PsiInitFromRAI-style factory (deconvolution/init/PsiInitFromRAIFactory.java:28-48). */
int mvr_unet_load(mvr_unet** out, int dev, const char* path);
/* host only (no GPU): the weights file's parameter count (returned), parameters (if params != NULL and
 * cap >= count) and views; -1 on a malformed file */
long long mvr_unet_read_weights(const char* path, float* params, long long cap, int* views);
/* number of FP32 parameters; get/set use one flat array in layer order (per conv layer the weights
 * [27][cout][cin] then the bias [cout]; last the 1x1 head: 16 weights and its bias) */
long long mvr_unet_param_count(mvr_unet* u);
int mvr_unet_get_params(mvr_unet* u, float* out);
int mvr_unet_set_params(mvr_unet* u, const float* in);
/* gradients of the last mvr_unet_train_step, same layout as the parameters */
int mvr_unet_get_grads(mvr_unet* u, float* out);
/* copy one training volume (V views, V weights, target psi, dims n) to the device; returns its index */
int mvr_unet_add_volume(mvr_unet* u, const float* const* imgs, const float* const* weights, const float* target,
                        const int n[3]);
int mvr_unet_clear_volumes(mvr_unet* u);
/* one step on batch patches of patch^3 voxels (volume volIdx[b], origin origins[3b..3b+2] as {x, y, z}):
 * forward, mean squared error to the normalised target in *loss, backward; Adam step when update != 0 */
int mvr_unet_train_step(mvr_unet* u, const int* volIdx, const int* origins, int batch, int patch, float lr,
                        int update, double* loss);
/* psiOut[n] = max(minValue, network estimate) for the given views and weights */
int mvr_unet_infer(mvr_unet* u, const float* const* imgs, const float* const* weights, const int n[3],
                   float minValue, float* psiOut);
/* Inference of volumes above `voxels` (padded to multiples of 4; default and any value <= 0: 2^28) runs in
 * z slabs that each carry mvr_unet_inference_halo() planes of context on either side; the result equals
 * the single-window result.  Bounds the network's activation memory beside a resident deconvolution
 * (about 285 bytes per window voxel).  No counterpart in the reference (its PsiInit has no network). */
int mvr_unet_set_inference_window(mvr_unet* u, long long voxels);
int mvr_unet_inference_halo(void);
void mvr_unet_destroy(mvr_unet* u);
/* mvr_decon_init with the network's estimate as psi (and as the psi restored by mvr_decon_reset_psi);
 * on an initialised handle only psi is replaced (PSFs, spectra, max[] and avg are kept) */
int mvr_decon_init_learned(mvr_decon* d, mvr_unet* net);

const char* mvr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
