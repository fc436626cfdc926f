/* C ABI of libcuhc.so, the MI355X (gfx950) build of cuhc's clustering path.
 *
 * The first eight functions are the drop-in boundary: the same names,
 * signatures, argument meanings, ID tables and status codes as the reference
 * library's extern "C" interface (reference src/capi.cu). Host pointers in and
 * out, blocking calls, one GPU (device 0). The functions after them have no
 * counterpart in the reference; each says what it is for.
 *
 * Status codes: 0 = success; a positive value is a hipError_t (2 = out of
 * memory); -1 invalid metric, -2 invalid shape, -3 invalid linkage,
 * -4 invalid algorithm (or core method) for the request, -5 disconnected (some
 * points have no finite distance to the rest), -6 non-finite merge heights.
 * cuhc_linkage, cuhc_ward_rnn_with_options, cuhc_hdbscan_tree and their
 * *_device forms return -6 before any distance is computed when X holds NaN or
 * infinity (the reference leaves that check to its Python layer).
 *
 * ID tables. metric: 0 euclidean, 1 sqeuclidean, 2 manhattan, 3 chebyshev,
 * 4 minkowski (metricParam = p), 5 cosine. linkage: 0 single, 1 complete,
 * 2 average, 3 ward. algorithm: 0 dense, 1 mst (single only), 2 nn_chain and
 * 3 rnn (ward + euclidean only). core method: 0 auto, 1 fused (k <= 64),
 * 2 radix, 3 learned radius (an extension; same values). prune mode: 0 auto,
 * 1 always, 2 never.
 */
#ifndef CUHC_H
#define CUHC_H

#ifdef __cplusplus
extern "C" {
#endif

/* distances (nA x nB, row-major) = metric(A[i], B[j]); B == NULL means B = A.
 * Replaces reference src/capi.cu:138-173 (cuhc_pairwise_distances). */
int cuhc_pairwise_distances(const float *A, int nA, const float *B, int nB, int d, int metricId,
                            float metricParam, float *distances);

/* Full merge tree in scipy's linkage format: tree receives (n - 1) x 4
 * doubles {id1, id2, height, size}. Replaces reference src/capi.cu:179-257. */
int cuhc_linkage(const float *X, int n, int d, int metricId, float metricParam, int linkageId,
                 int algorithmId, double *tree);

/* Ward through RNN rounds with explicit hand-over settings (test hook).
 * Replaces reference src/capi.cu:261-279. */
int cuhc_ward_rnn_with_options(const float *X, int n, int d, float minProgress, int denseBelow,
                               int allowDense, double *tree);

/* HDBSCAN's single-linkage tree of mutual reachability; core (n floats, may be
 * NULL) receives the core distances. Replaces reference src/capi.cu:285-309. */
int cuhc_hdbscan_tree(const float *X, int n, int d, int metricId, float metricParam,
                      int minSamples, float alpha, int coreMethod, double *tree, float *core);

/* Condensed tree from an (n - 1) x 4 single-linkage tree: writes up to
 * capacity rows and returns the row count, or -(rows needed).
 * Replaces reference src/capi.cu:314-327. */
long long cuhc_condense_tree(const double *tree, long long n, long long minClusterSize,
                             long long capacity, long long *parent, long long *child,
                             double *lambda, long long *size);

/* Spatial pruning mode for mst, rnn and HDBSCAN (results never change).
 * Replaces reference src/capi.cu:332-337. */
int cuhc_set_pruning(int mode);

/* Tiles computed / total by pruned passes since the last call, then resets.
 * Replaces reference src/capi.cu:341-345. */
void cuhc_prune_stats(unsigned long long *computed, unsigned long long *total);

/* Message for a status code. Replaces reference src/capi.cu:347-364. */
const char *cuhc_error_string(int code);

/* ---- Extensions (no counterpart in the reference) ---------------------- */

/* HDBSCAN cluster extraction on the host, the steps after cuhc_condense_tree
 * (the reference runs them in NumPy, python/cuhc/_hdbscan.py:148-350, with
 * the results of scikit-learn's _tree.pyx). The condensed tree is given as its
 * four columns of `rows` rows; root = min(parent), top = max(parent).
 * cuhc_cluster_stability: stability[c - root] for clusters root .. top.
 * cuhc_select_clusters: method 0 excess of mass, 1 leaf, without the epsilon
 * rule (the caller applies it); stability (as above) is updated as excess of
 * mass does; selected[c - root] = 1 for chosen clusters; maxClusterSize <= 0
 * means no limit. cuhc_label_points: labels[p] for points 0 .. root - 1 (-1 =
 * noise, else the rank of the point's cluster among the selected ids).
 * cuhc_membership_probabilities: probabilities[p] for those labels. Return 0
 * or a status. */
int cuhc_cluster_stability(const long long *parent, const long long *child, const double *lambda,
                           const long long *size, long long rows, double *stability);
int cuhc_select_clusters(const long long *parent, const long long *child, const double *lambda,
                         const long long *size, long long rows, double *stability, int method,
                         int allowSingleCluster, long long maxClusterSize,
                         unsigned char *selected);
int cuhc_label_points(const long long *parent, const long long *child, const double *lambda,
                      const long long *size, long long rows, const unsigned char *selected,
                      int allowSingleCluster, double epsilon, long long *labels);
int cuhc_membership_probabilities(const long long *parent, const long long *child,
                                  const double *lambda, const long long *size, long long rows,
                                  const unsigned char *selected, const long long *labels,
                                  double *probabilities);

/* Precision of the single-linkage Borůvka passes: 0 auto (FP16 MFMA
 * candidates + certified FP32 recheck for euclidean/sqeuclidean, d >= 16),
 * 1 plain FP32 tiles, 2 mixed whenever the metric allows. The tree is
 * identical in every mode. For benchmarks and tests. */
int cuhc_set_precision(int mode);

/* cuhc_linkage with X already in device memory (n x d floats, row-major,
 * allocated in this process's HIP runtime, e.g. by cuhc_device_alloc). tree is
 * host memory. The call synchronises the device before reading X, so X may
 * have been written on any stream. For callers that keep points resident in
 * HBM (zero-copy input). */
int cuhc_linkage_device(const float *deviceX, int n, int d, int metricId, float metricParam,
                        int linkageId, int algorithmId, double *tree);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory while they run (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). cuhc_linkage_device for single linkage
 * (linkageId 0, algorithmId 1) whose device buffers all come from one fixed
 * 256 MiB array inside the library (one per GPU, held until the process
 * ends): it calls no hipMalloc, hipFree, pinned-memory or other allocating
 * routine. At most 65536 points of at most 64 features; larger inputs return
 * -7 before any work. Same tree, statuses and precision/pruning modes as
 * cuhc_linkage_device. Calls are serialised within the process; calls that
 * overlap on one GPU by other means race on the shared array. */
int cuhc_linkage_device_noalloc(const float *deviceX, int n, int d, int metricId,
                                float metricParam, int linkageId, int algorithmId, double *tree);

/* The library's own device primitives, exposed as test hooks (no
 * counterpart in the reference, whose scans are CUB's): exclusive prefix sum
 * of ints, and stable LSD radix sorts of (key, value) pairs by the low `bits`
 * bits of the keys (higher bits zero). Device arrays; inputs unchanged
 * (in == out allowed for the scan); synchronous. Return a status. */
int cuhc_exclusive_scan(const int *in, int *out, int count);
int cuhc_radix_sort_pairs_u32(const unsigned int *keysIn, unsigned int *keysOut,
                              const int *valuesIn, int *valuesOut, int count, int bits);
int cuhc_radix_sort_pairs_u64(const unsigned long long *keysIn, unsigned long long *keysOut,
                              const unsigned int *valuesIn, unsigned int *valuesOut, int count,
                              int bits);

/* cuhc_pairwise_distances with every array in device memory (A, B nullable
 * as in the host form, out nA x nB floats): no PCIe copy. The device is
 * synchronised before X is read. No counterpart in the reference. */
int cuhc_pairwise_distances_device(const float *A, int nA, const float *B, int nB, int d,
                                   int metricId, float metricParam, float *out);

/* Per-feature minimum and maximum of device points (n x d floats) into the
 * host arrays low[d] and high[d]: the overflow check of the Python surface
 * (the float32 distance bound) for inputs already in HBM. */
int cuhc_feature_ranges_device(const float *X, int n, int d, float *low, float *high);

/* cuhc_hdbscan_tree with X in device memory; tree and core are host memory. */
int cuhc_hdbscan_tree_device(const float *deviceX, int n, int d, int metricId, float metricParam,
                             int minSamples, float alpha, int coreMethod, double *tree,
                             float *core);

/* Device buffers on the library's HIP runtime, for callers that keep inputs
 * resident in HBM across calls (the *_device entry points). Return a status.
 * cuhc_device_alloc releases the library's cache of freed blocks and retries
 * when device memory is short. */
int cuhc_device_alloc(unsigned long long bytes, void **ptr);
int cuhc_device_free(void *ptr);
int cuhc_copy_to_device(void *dst, const void *src, unsigned long long bytes);
int cuhc_copy_to_host(void *dst, const void *src, unsigned long long bytes);
/* Device-to-device copy (a plain copy kernel of known size: the benchmark
 * uses it to scale the HBM byte counters). No counterpart in the reference. */
int cuhc_copy_device(void *dst, const void *src, unsigned long long bytes);

/* Matrix-core issue-rate probe: times `workgroups` workgroups of 256 threads
 * running `iters` steps of one instruction mix (shape 0: 32x32x16 FP16 MFMAs
 * accumulating in place; 1: 16x16x32 FP16 MFMAs accumulating in place; 2: the
 * candidate pass's loop on synthetic LDS operands, seeded accumulators and the
 * per-lane minimum epilogue; 3: that loop on 16x16x32 MFMAs; 4 and 5: loops
 * 2 and 3 with fewer row blocks per wave at four waves per SIMD; 6: loop 2
 * with per-column tests as well, its flops counted twice) and returns ms
 * and flops per launch (-2 for a shape outside 0-6). No memory
 * traffic; the benchmark uses it to state an attainable MFMA ceiling. No
 * counterpart in the reference. */
int cuhc_mfma_probe(int shape, int workgroups, int iters, float *msPerLaunch,
                    double *flopsPerLaunch);

/* Frees what the library keeps between calls: its cache of freed device
 * blocks, the pinned staging buffer and its HIP stream (recreated on the next
 * call). Also runs from an exit handler, before the HIP runtime's teardown. */
void cuhc_release(void);

/* Per-phase GPU timing with HIP events on the library stream (0 = off).
 * Phases: 0 FP32 distance tiles, 1 FP16 MFMA candidates, 2 FP32 recheck,
 * 3 core distances, 4 component bookkeeping, 5 dense loop, 6 RNN kernels,
 * 7 learned-radius network; 8-11 split the dense loop by the device clock
 * into pair selection, Lance-Williams update, row rescans and grid barriers;
 * 12 host tree assembly from the downloaded edges (host clock). */
int cuhc_set_timing(int on);
/* Fills up to count phases: total ms, launches, algorithmic units (pairs). */
void cuhc_read_timings(double *ms, long long *launches, long long *units, int count);
void cuhc_reset_timings(void);

/* Learned-radius core distances (cuhc_hdbscan_tree coreMethod 3): counters
 * since the last call (then resets): radius proposals, proposals that held
 * (k neighbours inside the radius, list not overflowing), exact fallbacks, and
 * in-radius candidates gathered, distance tiles computed by the gather; then
 * (nullable) Boruvka rows answered exactly from the kept candidate lists and
 * rows offered to them, summed over the pruned rounds. */
void cuhc_learned_stats(unsigned long long *proposals, unsigned long long *hits,
                        unsigned long long *fallbacks, unsigned long long *candidates,
                        unsigned long long *tiles, unsigned long long *settled,
                        unsigned long long *considered);
/* Adam steps used to train the radius network per call (default 300; 0 keeps
 * the untrained initialisation — a test hook that measures what training
 * gains). Returns -2 for negative values. */
int cuhc_set_learned_training(int steps);

/* The radius-proposal network (4 inputs, two tanh layers of 32, one output),
 * host arrays in and out, for tests and offline use. Parameters:
 * cuhc_mlp_param_count() floats, layout W1[32][4], b1, W2[32][32], b2, W3[32], b3.
 * Gradient of the mean of 0.5 (y - t)^2 over m samples; loss may be NULL. */
int cuhc_mlp_param_count(void);
int cuhc_mlp_gradient(const float *params, const float *X, const float *t, int m, float *grad,
                      float *loss);
int cuhc_mlp_forward(const float *params, const float *X, int m, float *y);
/* Full-batch Adam, `steps` updates of params (in place). */
int cuhc_mlp_train(float *params, const float *X, const float *t, int m, int steps, float lr);

/* Counters of the mixed-precision pass since the last call (then resets):
 * candidate tiles recorded, tiles rechecked in FP32, pairs rechecked,
 * rows whose candidate list overflowed (rechecked in full). */
void cuhc_mixed_stats(unsigned long long *recorded, unsigned long long *rechecked,
                      unsigned long long *pairs, unsigned long long *overflows);

#ifdef __cplusplus
}
#endif

#endif /* CUHC_H */
