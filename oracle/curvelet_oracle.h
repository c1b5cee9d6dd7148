/* Test infrastructure only.
 *
 * CPU restatement of cpu_curvelet's curvelet-construction path, used by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker.  The product path (curvelet_construction_on_gpu_amd) never links,
 * loads or calls it.
 *
 * Parity pins (see tests/test_oracle_*.py): byte-exact chain/info text and
 * full-precision arrays against the reference built in place
 * (oracle/_ref, golden fixtures under tests/golden/), brute-force neighbour
 * search, closed-form bundle special cases and chain invariants.
 */
#ifndef CURVELET_ORACLE_H
#define CURVELET_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters after the casts of cpu_curvelet/main.cpp:22-30 (dt in radians). */
typedef struct orc_params {
    double nrad;
    double dx;
    double dt;
    double max_k;
    double sx;
    double st;
    unsigned group_max_sz;
    unsigned max_candidates;
    unsigned neighbor_radius; /* cells; main.cpp:48 passes 3 */
} orc_params;

/* dt = (dt_deg / 180) * pi, main.cpp:24. */
double orc_dt_from_deg(double dt_deg);

void *orc_create(void);
void orc_destroy(void *h);

/* build_neighbor_csr_graph (preprocess.hpp:186-274). Returns 0 on success. */
int orc_build_csr(void *h, int n, int sz, const double *edges, const orc_params *p, int nthreads);
void orc_csr_counts(void *h, int *pairs, unsigned *max_deg, unsigned *truncated_anchors);
void orc_get_csr(void *h, int *offsets, int *ids, double *dist2);

/* form_pairwise_bundles valid-pair count (cpu_curvelet.hpp:305-361). */
unsigned orc_valid_pairs(void *h, const orc_params *p, int nthreads);

/* build_curvelets_greedy (cpu_curvelet.hpp:363-634). Returns the curvelet count. */
unsigned orc_greedy(void *h, const orc_params *p, int nthreads);
unsigned orc_chain_width(void *h);
/* chains: C x (G+1) 1-based ids; info: C x 10 row-major. */
void orc_get_curvelets(void *h, unsigned *chains, double *info);

/* One bundle (compute_curve_bundle, cpu_curvelet.hpp:636-819); returns the
 * hypothesis flag, writes K = Kx*Kt cells. */
int orc_pair_bundle(const orc_params *p, double ax, double ay, double at, double nx, double ny, double nt,
                    double *bmin, double *bmax);
int orc_bundle_cells(const orc_params *p, unsigned *kx, unsigned *kt);

/* Direction filter value dot(angle_from_pt_to_pt(a, n), theta_a) as the
 * forward (run 0) or backward (run 1) pass of the reference evaluates it. */
double orc_dir_dot(int run, double ax, double ay, double at, double nx, double ny);

/* Per-anchor detail for GPU-stage parity: for each CSR neighbour r of anchor a,
 * dir[r] = bit 0 forward run, bit 1 backward run (both only if the two
 * evaluations disagree at a zero dot product), hyp[r] = bundle
 * valid or close, bmin/bmax = deg x K cells. */
int orc_anchor_detail(void *h, const orc_params *p, int anchor, unsigned char *dir, unsigned char *hyp,
                      double *bmin, double *bmax);

#ifdef __cplusplus
}
#endif

#endif
