/* curvelet_gpu.h - C ABI of the MI355X curvelet-construction library
 * (libcurvelet_gpu.so).
 *
 * The reference (cpu_curvelet) has no FFI: its boundary is the CURVELET command
 * line plus the in-process C++ API that run_curvelet<T> calls
 * (cpu_curvelet/main.cpp:19-96).  Each entry point below replaces one piece of
 * that API; the comment names the reference interface it stands in for.
 * Every function returns a cvg_status (0 = success), never aborts and never
 * lets a C++ exception cross the boundary; cvg_last_error() describes the last
 * failure on a context.  Plain pointers and sizes only.
 *
 * Threading: one context per host thread.  A context owns one HIP stream and
 * device-resident pools (grow-only) for the spatial hash, the neighbour graph,
 * the curvelet staging rows and the outputs.  Results never depend on stream,
 * batch composition or launch shape.
 */
#ifndef CURVELET_GPU_H
#define CURVELET_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference returns bool and prints to std::cerr
 * (data_io.hpp:24-75, preprocess.hpp:197, main.cpp:38-59); these split that
 * false into causes. */
typedef enum cvg_status {
    CVG_OK = 0,
    CVG_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, bad parameter */
    CVG_ERR_BAD_INPUT = 2,        /* malformed edge file or non-finite / out-of-range edge values */
    CVG_ERR_IO = 3,               /* file missing or unwritable */
    CVG_ERR_OUT_OF_MEMORY = 4,    /* host or device allocation failed */
    CVG_ERR_DEVICE = 5,           /* HIP runtime error */
    CVG_ERR_UNSUPPORTED = 6,      /* parameters outside what the kernels are built for */
    CVG_ERR_STATE = 7             /* call order: e.g. curvelets before the neighbour graph */
} cvg_status;

/* CurveletParams, cpu_curvelet/param_settings.hpp:9-24, same fields and
 * defaults.  dt_deg stays in degrees and is converted exactly as
 * main.cpp:24 does: dt = (dt_deg / 180) * pi.  gap, curvelet_style and
 * out_type are accepted and ignored, as in the reference. */
typedef struct cvg_params {
    double nrad;
    double gap;
    double dx;
    double dt_deg;
    double max_k;
    unsigned curvelet_style;
    unsigned group_max_sz;
    unsigned out_type;
    double sx;
    double st;
    int edge_data_sz;
    unsigned max_candidates;
} cvg_params;

/* Scalar members of CPUNeighborGraph<T> (preprocess.hpp:17-31) plus the
 * library's own diagnostics. */
typedef struct cvg_graph_info {
    int num_edges;
    int total_neighbor_pairs;
    unsigned max_num_of_neighbors;
    /* anchors whose list reached max_candidates (the reference's warning count, preprocess.hpp:224,264-268) */
    unsigned truncated_anchors;
    /* anchors whose neighbour order depends on the within-cell order that
     * libstdc++'s unstable std::sort gives the reference's spatial index
     * (exact distance ties inside one pixel cell, a max_candidates cut inside
     * one cell, or ties in a list of more than 16 that holds two candidates
     * of one cell).  When there are any, the library replays that sort on the
     * host and builds the graph again in the reference's order, so the
     * results stay exact; the count is diagnostic. */
    unsigned order_ambiguous_anchors;
} cvg_graph_info;

/* Timings are device time measured with HIP events on the context's stream. */
typedef struct cvg_run_stats {
    unsigned valid_pairs;      /* form_pairwise_bundles count (main.cpp:61) */
    unsigned num_curvelets;    /* out_h (main.cpp:68) */
    unsigned chain_width;      /* out_w = group_max_sz + 1 (main.cpp:69) */
    int total_neighbor_pairs;
    unsigned max_num_of_neighbors;
    double ms_graph;           /* spatial hash + CSR neighbour graph */
    double ms_curvelets;       /* bundles + chain growth + dedup + record */
    double ms_total;
    double ms_fused_kernel;    /* the bundle + growth kernel launches alone */
    unsigned long long intersections; /* bundle intersections tested during growth */
    unsigned order_ambiguous_anchors;
    unsigned lds_bytes_per_workgroup; /* the device limit the kernels were sized for */
} cvg_run_stats;

typedef struct cvg_context cvg_context;

/* ---- library / context ------------------------------------------------ */

/* CurveletParams{} default member initialisers (param_settings.hpp:10-23). */
void cvg_params_default(cvg_params *params);
const char *cvg_status_string(int status);
/* Bind a context to HIP device `device` with its own stream. */
int cvg_create(int device, cvg_context **out_ctx);
void cvg_destroy(cvg_context *ctx);
const char *cvg_last_error(const cvg_context *ctx);

/* ---- edge-file codec (data_io.hpp) ------------------------------------ */

/* read_TO_edges_from_file<double> (data_io.hpp:20-78) minus the
 * "../test_files/" prefix (edge_input_path, data_io.hpp:10-13), which the CLI
 * adds: exactly values_per_edge numbers per line, blank / short / long lines
 * and empty files rejected with the reference's messages (written to err).
 * *edges is allocated by the library (release with cvg_free). */
int cvg_read_edge_file(const char *path, int values_per_edge, double **edges, int *num_edges, char *err,
                       size_t err_len);
/* read_TO_edges_from_file<float> (data_io.hpp:20-78 on T = float, the --float
 * mode of main.cpp:110): the same grammar and messages, each value converted
 * from its decimal text directly to float (istream >> float). */
int cvg_read_edge_file_f32(const char *path, int values_per_edge, float **edges, int *num_edges, char *err,
                           size_t err_len);
void cvg_free(void *ptr);
/* write_int_array_to_file / write_double_array_to_file (data_io.hpp:80-133):
 * tab after every value, newline per row, doubles in default ostream format
 * (6 significant digits).  Writes exactly `path`; the CLI adds "../outputs/". */
int cvg_write_chain_file(const char *path, const unsigned *chains, unsigned rows, unsigned width);
int cvg_write_info_file(const char *path, const double *info, unsigned rows, unsigned width);

/* ---- stage-wise mirror of the reference's in-process API -------------- */

/* build_neighbor_csr_graph(num_edges, sz_edge_data, edges, rad, neighbor_radius,
 * max_candidates, nthreads, graph) (preprocess.hpp:186-274).  `edges` is a
 * host array, row-major num_edges x sz_edge_data (sz_edge_data >= 4). */
int cvg_build_neighbor_csr_graph(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges,
                                 double rad, unsigned neighbor_radius, unsigned max_candidates);
/* build_neighbor_csr_graph<float>: the same on a float edge array; the
 * context then runs the float instantiation (run_curvelet<float>,
 * main.cpp:110) in the stage-wise calls below: float distances, bundles in
 * compute_curve_bundle<float>'s arithmetic, info in fill_curvelet_info<float>'s. */
int cvg_build_neighbor_csr_graph_f32(cvg_context *ctx, int num_edges, int sz_edge_data, const float *edges,
                                     double rad, unsigned neighbor_radius, unsigned max_candidates);
int cvg_get_graph_info(cvg_context *ctx, cvg_graph_info *info);
/* neighbor_offsets (num_edges + 1), neighbor_ids and neighbor_dist2
 * (total_neighbor_pairs each) of CPUNeighborGraph<T>. */
int cvg_get_csr(cvg_context *ctx, int *offsets, int *ids, double *dist2);

/* CurveletCPU<T>::form_pairwise_bundles (cpu_curvelet.hpp:305-361): the
 * valid-pair count.  Runs the fused bundle + growth pass and keeps its
 * results, so a following cvg_build_curvelets_greedy with the same params
 * does not recompute. */
int cvg_form_pairwise_bundles(cvg_context *ctx, const cvg_params *params, unsigned *valid_pairs);
/* CurveletCPU<T>::build_curvelets_greedy (cpu_curvelet.hpp:363-634) with
 * compact_curvelet_output (:282-303). */
int cvg_build_curvelets_greedy(cvg_context *ctx, const cvg_params *params, unsigned *num_curvelets);
/* num_curvelets() and chain_width() (cpu_curvelet.hpp:95-96). */
int cvg_curvelet_shape(cvg_context *ctx, unsigned *num_curvelets, unsigned *chain_width);
/* _edge_chain_final (num_curvelets x chain_width, 1-based ids, row-major) and
 * _curvelet_info transposed to row-major num_curvelets x 10 as
 * main.cpp:81-88 does. */
int cvg_get_curvelets(cvg_context *ctx, unsigned *chains, double *info);

/* ---- whole path --------------------------------------------------------- */

/* The timed span of run_curvelet<double> (main.cpp:43-66): edges in host
 * memory -> graph -> bundles -> chains and info on the device.  Fetch the
 * results with cvg_get_curvelets.  With stats == NULL this and the other
 * run calls return once the work is queued (after the graph stage's two host
 * reads, so the host edge arrays are no longer needed); the results and
 * counts are complete at their first query (cvg_curvelet_shape,
 * cvg_get_curvelets, cvg_device_curvelets, ...).  With stats the call waits
 * and fills them. */
int cvg_run(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges, const cvg_params *params,
            cvg_run_stats *stats);

/* run_curvelet<double> from host edges to host results (main.cpp:43-88: the
 * timed span plus the copy of _edge_chain_final and _curvelet_info into the
 * output arrays) in one call: chains (max_rows x (group_max_sz + 1),
 * 1-based) and info (max_rows x 10, row-major) in host memory (page-locked
 * ones from cvg_host_alloc take one DMA each), *num_rows the row count.
 * Returns when the rows are in the outputs.  More rows than max_rows fails
 * with CVG_ERR_INVALID_ARGUMENT after writing the first max_rows (*num_rows
 * still says how many there are).  Rows past *num_rows (up to max_rows) may
 * be overwritten: page-locked outputs receive as many rows as the context's
 * previous call returned as soon as the record ends, before the row count
 * is read back. */
int cvg_run_to_host(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges, const cvg_params *params,
                    unsigned *chains, double *info, unsigned max_rows, unsigned *num_rows, cvg_run_stats *stats);
/* run_curvelet<float> (main.cpp:110 with --float): the whole path on a float
 * edge array (row-major num_edges x sz_edge_data), in the reference's float
 * arithmetic; results as for cvg_run (chains, and info widened to double). */
int cvg_run_f32(cvg_context *ctx, int num_edges, int sz_edge_data, const float *edges, const cvg_params *params,
                cvg_run_stats *stats);
/* Same, with the edge array already in device memory (for instance from
 * cvg_upload_edges).  No host<->device copy of edges or results. */
int cvg_run_device(cvg_context *ctx, int num_edges, int sz_edge_data, const double *device_edges,
                   const cvg_params *params, cvg_run_stats *stats);
/* Size of the static device scratch array behind cvg_run_device_noalloc, in
 * MiB.  Every device the library runs on holds one copy of it until the
 * process ends. */
#define CVG_NOALLOC_SCRATCH_MIB 64

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): cvg_run_device
 * with every intermediate buffer (cell grid, neighbour graph, pair records and
 * sines, staging rows, counters) carved from the library's static device
 * scratch array of CVG_NOALLOC_SCRATCH_MIB MiB instead of the context's pools,
 * and the results written straight to caller device buffers: device_chains
 * (max_rows x (group_max_sz + 1) unsigned, 1-based ids) and device_info
 * (max_rows x 10 double, row-major).  Calls no hipMalloc, hipFree or other
 * allocating routine.  Waits for the run; *num_curvelets receives the row
 * count (also when it exceeds max_rows, which fails with
 * CVG_ERR_INVALID_ARGUMENT and leaves the rows beyond max_rows unwritten).  A
 * map whose buffers do not fit the array fails with CVG_ERR_UNSUPPORTED (about
 * 10k edges fit at the default parameters and G = 7).  The array is shared by every
 * context on a device and keeps the last call's contents (the run clears what
 * must start at zero): calls that overlap on one device race.  The context's
 * own results of earlier runs are no longer available afterwards. */
int cvg_run_device_noalloc(cvg_context *ctx, int num_edges, int sz_edge_data, const double *device_edges,
                           const cvg_params *params, unsigned *device_chains, double *device_info, unsigned max_rows,
                           unsigned *num_curvelets, unsigned *valid_pairs);
/* Copy a host edge array into a context-owned device buffer (valid until the
 * next upload or cvg_destroy). */
int cvg_upload_edges(cvg_context *ctx, int num_edges, int sz_edge_data, const double *edges,
                     const double **device_edges);
/* Device pointers to the last run's outputs (valid until the next run). */
int cvg_device_curvelets(cvg_context *ctx, const unsigned **device_chains, const double **device_info);
/* Page-locked host memory: edge arrays and result buffers in it transfer by
 * DMA at full link speed (the served form keeps pools of it). */
int cvg_host_alloc(size_t bytes, void **ptr);
void cvg_host_free(void *ptr);
/* The context's HIP stream (hipStream_t), for callers that order their own work after it. */
int cvg_get_stream(cvg_context *ctx, void **hip_stream);
int cvg_synchronize(cvg_context *ctx);
/* Durations in ms (HIP events on the context stream around the fused kernel's
 * launches, k_curvelets) of the most recent timed runs, oldest first: up to
 * min(max, 64) values, *count = how many.  Waits for the stream.  A run is
 * timed when it returns stats, or while cvg_set_event_timing is on.  No
 * counterpart in the reference: it times the device work of runs made
 * without stats. */
int cvg_recent_fused_ms(cvg_context *ctx, int max, float *ms, int *count);
/* on != 0: runs made without stats record the HIP timing events as well (the
 * stage and fused-kernel times that cvg_recent_fused_ms reads).  Off at
 * creation: each timing event is a marker that holds the stream for ~6 us,
 * seven of them per run.  Runs that return stats always record them.  No
 * counterpart in the reference (its timers are host-side). */
int cvg_set_event_timing(cvg_context *ctx, int on);

/* ---- batched form ---------------------------------------------------------
 * Several maps in one pass: every kernel of the path runs once over all of
 * them (each map's cells occupy their own region of one cell grid, so no
 * neighbour window crosses maps; anchors of all maps share the fused kernel's
 * launches).  Each map's chains and info are bit-identical to what cvg_run
 * gives for that map alone; chain ids are 1-based within their map.  The
 * reference has no batch interface: these stand in for a loop of
 * run_curvelet<double> calls (cpu_curvelet/main.cpp:20-96), one per map. */
int cvg_run_batch(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data,
                  const double *const *edges, const cvg_params *params, cvg_run_stats *stats);
/* Same, with the maps already concatenated (map-major) in device memory. */
int cvg_run_batch_device(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data,
                         const double *device_edges, const cvg_params *params, cvg_run_stats *stats);
/* One map's output shape (num_curvelets(), chain_width()) and rows. */
int cvg_batch_map_shape(cvg_context *ctx, int map, unsigned *num_curvelets, unsigned *chain_width);
int cvg_get_batch_map_curvelets(cvg_context *ctx, int map, unsigned *chains, double *info);

/* ---- launch tuning ---------------------------------------------------------
 * The launch shape of a context's kernels.  Results never depend on it; only
 * the time does.  The reference has no counterpart (it is a CPU loop); these
 * serve the batched and served forms, whose map mixes differ from the
 * defaults' 1920x1080 tuning.  The environment variables CVG_COOP_MIN_CAP,
 * CVG_COOP_WAVES and CVG_FILL_LANES set the initial values of a new context. */
typedef struct cvg_tuning {
    int coop_min_key;  /* anchor buckets whose key (max(forward, backward) + neither
                          neighbours) is at least this run one anchor per workgroup of
                          coop_waves waves; lighter ones one anchor per wave (>= 1) */
    int coop_waves;    /* 1, 2 or 4 */
    int window_lanes;  /* lanes per anchor in the neighbour-window scans: 8, 16, 32 or 64 */
    int grid_pct;      /* grid of the light buckets' curvelet launches as a percentage of
                          their resident capacity, 5..100; 0 = automatic (10 plus one
                          percent per 4,000 edges in the run, at most 100) */
} cvg_tuning;
typedef struct cvg_tuning_trial {
    cvg_tuning tuning;
    double ms;         /* mean wall time per run of the whole path, results in HBM */
} cvg_tuning_trial;
int cvg_get_tuning(cvg_context *ctx, cvg_tuning *tuning);
int cvg_set_tuning(cvg_context *ctx, const cvg_tuning *tuning);
/* Autotuner for the batched form: times cvg_run_batch_device on the caller's
 * maps under candidate tunings (coordinate descent over the four fields,
 * one warm-up and `reps` timed runs each) and keeps the fastest on the
 * context.  best/best_ms receive the result; trials (capacity max_trials, may
 * be null) every candidate tried, in order; num_trials their count.  The last
 * run's results stay on the context as after cvg_run_batch_device. */
int cvg_autotune_batch(cvg_context *ctx, int num_maps, const int *num_edges, int sz_edge_data,
                       const double *device_edges, const cvg_params *params, int reps, cvg_tuning *best,
                       double *best_ms, cvg_tuning_trial *trials, int max_trials, int *num_trials);

/* ---- served form ----------------------------------------------------------
 * A request loop with continuous batching.  Requests (one edge map each) queue
 * as they arrive; each of `workers` threads owns a context (streams and device
 * pools resident across batches) and, whenever free, admits the oldest
 * pending requests while their estimated HBM footprint (num_edges x
 * bytes_per_edge) fits hbm_budget_bytes, at most max_batch_maps, and runs them
 * as one batch.  Each result equals cvg_run on that map alone.  The reference
 * has no serving layer: this is the loop a service would run around
 * run_curvelet<double> (cpu_curvelet/main.cpp:20-96). */
typedef struct cvg_server cvg_server;
typedef struct cvg_server_config {
    int workers;             /* contexts (and host threads) running batches concurrently, 1..8 */
    int max_batch_maps;      /* maps per batch; 1 gives the plain one-map-at-a-time loop */
    double bytes_per_edge;   /* HBM footprint estimate per edge (edge arrays + neighbours x per-pair buffers) */
    double hbm_budget_bytes; /* per batch */
} cvg_server_config;
typedef struct cvg_server_result {
    int status;
    unsigned num_curvelets;
    unsigned chain_width;
    int num_edges;
    int batch_maps;          /* maps in the batch that carried this request */
    double queue_ms;         /* submit -> batch start */
    double latency_ms;       /* submit -> results in host memory */
} cvg_server_result;

void cvg_server_config_default(cvg_server_config *config);
int cvg_server_create(int device, const cvg_params *params, const cvg_server_config *config, cvg_server **out);
/* Copies the edges; returns at once with the request's id. */
int cvg_server_submit(cvg_server *server, int num_edges, int sz_edge_data, const double *edges,
                      unsigned long long *request_id);
/* Blocks until the request is done; fills its shape and timings. */
int cvg_server_wait(cvg_server *server, unsigned long long request_id, cvg_server_result *result);
/* Copies a finished request's chains (num_curvelets x chain_width, 1-based ids)
 * and info (num_curvelets x 10) and releases it. */
int cvg_server_fetch(cvg_server *server, unsigned long long request_id, unsigned *chains, double *info);
/* Zero-copy access to a finished request's rows in the server's page-locked
 * result block; valid until cvg_server_release. */
int cvg_server_view(cvg_server *server, unsigned long long request_id, const unsigned **chains, const double **info);
int cvg_server_release(cvg_server *server, unsigned long long request_id);
int cvg_server_counters(cvg_server *server, unsigned long long *batches, unsigned long long *maps,
                        unsigned long long *edges);
const char *cvg_server_last_error(cvg_server *server);
/* Finishes the queued requests, then stops the workers. */
void cvg_server_destroy(cvg_server *server);

/* ---- synthetic third-order edge maps ------------------------------------ */

/* Seeded generator of edge maps in the test_files format (x, y, theta,
 * strength).  No counterpart in the reference; used for benchmarks and tests. */
typedef struct cvg_synth_config {
    int width;
    int height;
    int num_curves;
    double radius_min;        /* circles / arcs */
    double radius_max;
    int radius_log_uniform;   /* 1: radius = 1/curvature with curvature log-uniform */
    double segment_min;       /* line segments */
    double segment_max;
    double bezier_extent;     /* control points within this box (<= 0: the whole image) */
    double arc_max_length;    /* cap on one arc's sampled length (<= 0: none) */
    double clutter_fraction;  /* share of all edges that are clutter */
    int junction_patches;     /* T-junction patches */
    int curves_per_patch;
    double patch_size;
    double pos_sigma;         /* px */
    double theta_sigma_deg;
} cvg_synth_config;

/* Named presets: "C1" (320x240, 40 curves, ~6k edges), "C2" (480x320,
 * 150 curves, ~20k), "C3" (1920x1080, 1500 curves, 40 T-junction patches of
 * 20 crossing segments, 30% clutter, ~207k). */
int cvg_synth_preset(const char *name, cvg_synth_config *config);
/* Writes up to `capacity` edges (4 doubles each) to `edges` (may be NULL to
 * only count) and the total count to *num_edges.  Deterministic in (config, seed). */
int cvg_synth_generate(const cvg_synth_config *config, unsigned long long seed, double *edges, int capacity,
                       int *num_edges);

/* ---- diagnostics ------------------------------------------------------- */

/* Per-neighbour detail of one anchor of the current graph, computed on the
 * device by the same device functions the fused kernel uses: dir[r] = 1
 * (forward run), 2 (backward run), 0 (neither); hyp[r] = bundle valid or
 * close; bmin/bmax = deg x K cells in the reference's cell order
 * (i * Kt + j).  Mirrors compute_curve_bundle (cpu_curvelet.hpp:636-819) and
 * the direction filter (:459-461). */
int cvg_debug_anchor_bundles(cvg_context *ctx, const cvg_params *params, int anchor, unsigned char *dir,
                             unsigned char *hyp, double *bmin, double *bmax, int *deg, int *cells);
/* The within-cell order the library replays for order-ambiguous maps: for
 * each edge, its position inside its pixel cell after the reference's
 * std::sort of the edge ids by packed cell key (build_spatial_index_cpu,
 * preprocess.hpp:78-92), restated as libstdc++'s introsort.  Host only. */
int cvg_debug_cell_order(int num_edges, const double *x, const double *y, int *rank);
/* Evaluate the library's correctly rounded math on n arguments.
 * op: 0 sin(x), 1 cos(x), 2 atan2(y, x), 3 sqrt(x), 4 To2Pi(x), 5 fmod(x, y),
 * 6 and 7 the sine and cosine of the joint sincos(x), 8 sin(x - y) as the
 * per-pair transport sines compute it (subtraction formula on double-double
 * sin/cos of x and y, else the direct evaluation), 9 whether that formula
 * settled the rounding (1.0) or not (0.0), 10-13 the double-double sin(x)
 * high and low parts and cos(x) high and low parts; 14-18 the binary32
 * functions of the float mode on float(x) (and float(y)): sinf, cosf,
 * atan2f(y, x), To2Pi, sqrtf.
 * ctx == NULL evaluates the host build of the same code; otherwise the device. */
int cvg_math_eval(cvg_context *ctx, int op, int n, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif

#endif
