/*
 * dcx.h -- C-ABI of the MI355X-native D-core ((k,l)-core) library.
 *
 * The drop-in boundary for the reference's decomposition and maintenance
 * path (llysdal/d-core-cuda).  Plain pointers and sizes only; every call
 * returns a dcx_status.  Host buffers are owned by the caller; device memory
 * and the HIP stream are owned by the library (a dcx_ctx per device).
 *
 * Types follow the reference (src/common.h:59-62): vertex ids and CSR
 * offsets are uint32, degrees and (k,l) values are int32.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef DCX_H
#define DCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dcx_status {
    DCX_OK = 0,
    DCX_ERR_INVALID_ARG = 1,  /* bad pointer, size or value */
    DCX_ERR_NO_DEVICE = 2,    /* no HIP device / ordinal out of range */
    DCX_ERR_HIP = 3,          /* a HIP runtime call failed */
    DCX_ERR_OUT_OF_MEMORY = 4,/* device or host allocation failed */
    DCX_ERR_CAPACITY = 5,     /* a bounded buffer or slack is exhausted */
    DCX_ERR_IO = 6,           /* file could not be opened or written */
    DCX_ERR_FORMAT = 7,       /* malformed file */
    DCX_ERR_NOT_FOUND = 8,    /* e.g. delete of an absent edge */
    DCX_ERR_STATE = 9,        /* call order violated (e.g. no decomposition yet) */
    DCX_ERR_BUDGET = 10,      /* request does not fit the memory budget */
    DCX_ERR_TIMEOUT = 11,     /* a bounded wait expired */
    DCX_ERR_INTERNAL = 12
} dcx_status;

/* Human-readable status name.  Replaces nothing: the reference has no error
 * reporting at all (no CUDA call is checked anywhere in main.cu). */
const char* dcx_status_str(dcx_status s);

/* ------------------------------------------------------------------ */
/* Host-side graph building (no GPU needed).                          */
/* ------------------------------------------------------------------ */

/* Seeded directed G(n,m): E distinct pairs, no self-loops, uniform endpoints,
 * sorted by (src,dst).  Replaces generateGraph (src/graph_gen.h:8,
 * src/graph_gen.cpp:9-34), which seeds from time() and writes a file. */
dcx_status dcx_gen_er(uint32_t V, uint64_t E, uint64_t seed, uint32_t* src, uint32_t* dst);

/* Seeded R-MAT: V = 2^scale, `samples` draws with quadrant probabilities
 * (a,b,c,1-a-b-c), self-loops and duplicates removed, sorted by (src,dst).
 * *E_out receives the edge count; returns DCX_ERR_CAPACITY if it exceeds
 * `cap` (pass cap = samples to be safe).  No reference counterpart
 * (the reference only loads SNAP files). */
dcx_status dcx_gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c,
                        uint64_t seed, uint32_t* src, uint32_t* dst, uint64_t cap,
                        uint64_t* E_out);

/* Seeded Chung-Lu digraph with power-law out/in weights (exponent gamma),
 * `samples` draws, deduplicated, sorted.  No reference counterpart. */
dcx_status dcx_gen_chunglu(uint32_t V, uint64_t samples, double gamma, uint64_t seed,
                           uint32_t* src, uint32_t* dst, uint64_t cap, uint64_t* E_out);

/* Seeded stream of update batches against an R-MAT graph of V = 2^scale
 * vertices with edges (src, dst): batch b has floor(exp(U ln max_batch))
 * edits (clipped to [1, max_batch]), each an insert with probability
 * p_insert (an R-MAT(a,b,c) sample not in the graph) or else a delete (a
 * uniformly chosen current edge); the graph evolves batch by batch.  The
 * edits of batch b are ins_pairs[2*ins_off[b] .. 2*ins_off[b+1]) and the
 * same for deletions.  With ins_pairs and del_pairs NULL only the offsets
 * are filled (sizing).  Replaces the one-edge samplers getRandomInsertEdge /
 * getRandomDeleteEdge (src/graph.cpp:110-144) for batched streams. */
dcx_status dcx_gen_update_stream(uint32_t V, uint64_t E, const uint32_t* src,
                                 const uint32_t* dst, uint32_t scale, double a, double b,
                                 double c, uint32_t nbatches, uint32_t max_batch,
                                 double p_insert, uint64_t seed, uint64_t* ins_off,
                                 uint64_t* del_off, uint32_t* ins_pairs, uint64_t ins_cap,
                                 uint32_t* del_pairs, uint64_t del_cap);

/* Edge list -> in/out CSR with `slack` free slots per vertex:
 * off[v+1]-off[v] = deg(v)+slack, neighbours in input order, free slots
 * 0xFFFFFFFF.  Arrays: off[V+1], deg[V], nbr[E + slack*V].
 * Replaces the CSR build of Graph::readFile (src/graph.cpp:180-243). */
dcx_status dcx_csr_build(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                         uint32_t slack, uint32_t* out_off, int32_t* out_deg, uint32_t* out_nbr,
                         uint32_t* in_off, int32_t* in_deg, uint32_t* in_nbr);

/* Text edge list "# V" + "s t" lines: count, then read.  `prune` = 0 keeps
 * every id (GraphCPU::readFile, src/graph_cpu.cpp:38-102); prune = 1 drops
 * vertices without edges and renumbers (Graph::readFile + pruneEmptyVertices,
 * src/graph.cpp:110-176, with V taken from the header instead of the
 * hard-coded 7414865 at src/graph.cpp:158).  Self-loops are dropped. */
dcx_status dcx_text_count(const char* path, uint32_t* V_header, uint64_t* lines);
dcx_status dcx_text_read(const char* path, int32_t prune, uint32_t* src, uint32_t* dst,
                         uint64_t cap, uint32_t* V_out, uint64_t* E_out);
/* Writes the same format (src/graph_gen.cpp:27-29, Graph::writeFile src/graph.cpp:252-259). */
dcx_status dcx_text_write(const char* path, uint32_t V, uint64_t E, const uint32_t* src,
                          const uint32_t* dst);

/* Binary CSR cache, byte-compatible with <file>-binary-g=N
 * (Graph::writeBinary / readBinary, src/graph.cpp:261-325). */
dcx_status dcx_binary_header(const char* path, uint32_t* V, uint32_t* E);
dcx_status dcx_binary_read(const char* path, uint32_t gap, uint32_t* out_off, int32_t* out_deg,
                           uint32_t* out_nbr, uint32_t* in_off, int32_t* in_deg,
                           uint32_t* in_nbr);
dcx_status dcx_binary_write(const char* path, uint32_t V, uint32_t E, uint32_t gap,
                            const uint32_t* out_off, const int32_t* out_deg,
                            const uint32_t* out_nbr, const uint32_t* in_off,
                            const int32_t* in_deg, const uint32_t* in_nbr);

/* .decomp result cache, byte-compatible (readDecompFile / writeDecompFile,
 * src/utils.h:76-109): i32 kmax, i32 lmax, i32 kmaxes[V], (kmax+1) x i32[V]. */
dcx_status dcx_decomp_write(const char* path, uint32_t V, int32_t kmax, int32_t lmax,
                            const int32_t* kmaxes, const int32_t* rows);
dcx_status dcx_decomp_header(const char* path, int32_t* kmax, int32_t* lmax);
dcx_status dcx_decomp_read(const char* path, uint32_t V, int32_t* kmaxes, int32_t* rows);

/* Ragged per-vertex text result (writeDecompFileCPU, src/utils.h:110-124). */
dcx_status dcx_decompcpu_write(const char* path, uint32_t V, const int64_t* row_ptr,
                               const int32_t* values);

/* ------------------------------------------------------------------ */
/* Device context and device-resident graphs.                         */
/* ------------------------------------------------------------------ */

typedef struct dcx_ctx dcx_ctx;
typedef struct dcx_graph dcx_graph;

/* Exact kernel variants; every variant gives bit-identical results. */
typedef enum dcx_variant {
    DCX_VARIANT_AUTO = 0,    /* library choice (batched rows) */
    DCX_VARIANT_PEEL = 1,    /* plain: one k per launch, whole grid per pass */
    DCX_VARIANT_BATCHED = 2, /* all k-passes in one persistent launch */
    DCX_VARIANT_HINDEX = 3,  /* h-index fixpoint per k */
    /* pass log only: a batched row derived from the row above by h-index
     * refinement with per-vertex supports (not a dcx_decompose argument) */
    DCX_VARIANT_DERIVED = 4
} dcx_variant;

typedef struct dcx_config {
    int32_t device;          /* HIP device ordinal */
    int32_t variant;         /* default dcx_variant for dcx_decompose */
    /* Batched-rows launch shape (0 = library default; dcx_autotune picks
     * them per graph): */
    int32_t rows_slots;      /* workgroups of the launch (default 2 per CU) */
    int32_t team_max;        /* most workgroups sharing one k-pass (default 32) */
    int32_t shrink_at;       /* a team finishes a pass on one workgroup once
                                this few vertices are alive (default 4096) */
    int32_t trace_phases;    /* phase log for dcx_last_trace: 1 = grid passes
                                (kmax pass, plain rows); 2 + k = the team pass
                                of row k in the batched launch; 0 = off */
    uint64_t scratch_budget; /* max bytes of pass scratch (0 = no limit) */
    /* A cascade phase of a peel pass pulls (each alive vertex counts its
     * neighbours in the frontier) instead of pushing decrements from the
     * frontier when the alive vertices' lists and the alive list hold at
     * most pull_ratio times the frontier's edges: 0 = library default
     * (always push), -1 = always push.  Results are the same either way. */
    int32_t pull_ratio;
    /* Batched rows whose (k,0)-core has at most lds_members members run on
     * one workgroup with their degree counters in LDS: 0 = as many as the
     * device's LDS holds (at most 32767), -1 = none. */
    int32_t lds_members;
    /* Maintenance after an edge batch (dcx_update, the server):
     * 0 = local with a work budget (default): only the vertices the batch
     *     can change are visited -- h-index refinement from the old values
     *     after deletions, candidate regions with upper bounds after
     *     insertions -- and a batch that exceeds the budget (local_budget)
     *     or overflows a worklist is given up and recomputed instead;
     * 1 = recompute kmaxes and the rows k <= R;
     * 2 = local without the budget (still recomputed on overflow).
     * Results are identical. */
    int32_t maintenance;
    /* Budget of local maintenance: > 0 = list entries per edge that its
     * tasks may read (a deterministic choice of path); 0 = a quarter of the
     * graph's last full decomposition time (at least 0.2 ms), or 12 entries
     * per edge while no time is known. */
    int32_t local_budget;
    /* Batched rows: every chain_rows-th team row (from row 0) is peeled,
     * and the rows between are derived from the row above by h-index
     * refinement with per-vertex supports (only the cells that change are
     * visited); 0 = the default spacing, 1 or -1 = every row peeled.
     * Results are identical. */
    int32_t chain_rows;
    /* Batched rows, fixed rule: how an anchor row (one not derived from the
     * row above) is computed: 0 = peeled (default); 1 = derived from row
     * 0 (h-index refinement from row 0's values, which bound every row).
     * The cost model chooses per anchor.  Results are identical. */
    int32_t anchors;
    uint64_t reserved[1];
} dcx_config;

/* Timings of the last decomposition on a context, from HIP events on the
 * library's stream.  Mirrors the "kmax_ms lmax_ms total_ms" print of dcore
 * (main.cu:256). */
typedef struct dcx_timing {
    double kmax_ms;        /* in-core (kmax) pass kernel */
    double order_ms;       /* kmax-descending vertex order */
    double rows_ms;        /* all l-rows: clear, kernel launches, lmax readback */
    double rows_kernel_ms; /* the rows kernel launches alone */
    double total_ms;       /* kmax_ms + order_ms + rows_ms */
    int32_t rows_launches; /* kernel launches spent on the rows */
    int32_t rows_slots;    /* workgroups busy in the batched launch */
    int32_t rows_scheduler; /* batched plan: 0 fixed rule, 1 cost model */
    int32_t rows_split;    /* batched: rows k >= rows_split ran LDS-resident */
    float rows_teams_ms;   /* batched: the team launch alone */
    float rows_lds_ms;     /* batched: the LDS-resident launch alone (beside it) */
    float rows_first_ms;   /* batched: row 0's grid pass, run first to bound the others */
    float rows_chains_ms;  /* 0: rows derived from the row above run inside the
                              team launch, after their anchor's peel */
} dcx_timing;

dcx_status dcx_config_init(dcx_config* cfg);
dcx_status dcx_device_count(int32_t* n);
/* Creates a context on cfg->device: one HIP stream (default priority),
 * device scratch grown on demand.  Replaces the allocate*Memory family
 * (src/cuda_memory.h:5-106). */
dcx_status dcx_ctx_create(const dcx_config* cfg, dcx_ctx** out);
dcx_status dcx_ctx_destroy(dcx_ctx* ctx);
/* Detail of the last failing call on this context ("" if none). */
const char* dcx_ctx_last_error(const dcx_ctx* ctx);
/* Sets cfg.chain_rows, cfg.team_max and cfg.anchors of a context (which
 * batched team rows are anchors and which derived from the row above, the
 * widest team, and whether anchors are peeled or derived from row 0) for
 * its next decompositions; results are identical.  No reference
 * counterpart. */
dcx_status dcx_ctx_set_rows_rule(dcx_ctx* ctx, int32_t chain_rows, int32_t team_max,
                                 int32_t anchors);
dcx_status dcx_last_timing(const dcx_ctx* ctx, dcx_timing* t);

/* Uploads a CSR-with-slack graph (layout of dcx_csr_build).  Replaces
 * allocateDeviceGraphMemory + moveGraphToDevice (src/cuda_memory.h:5-41)
 * and the GraphGPU constructor (src/graph_gpu.cu:25-47). */
dcx_status dcx_graph_upload_csr(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* out_off,
                                const int32_t* out_deg, const uint32_t* out_nbr,
                                const uint32_t* in_off, const int32_t* in_deg,
                                const uint32_t* in_nbr, dcx_graph** out);
/* Builds the CSR (dcx_csr_build) and uploads it. */
dcx_status dcx_graph_upload_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                                  const uint32_t* dst, uint32_t slack, dcx_graph** out);
/* The same graph as dcx_graph_upload_edges (same layout, neighbours in
 * input order, duplicates kept), built on the GPU: degree count, offset
 * scan and a radix sort per side.  src/dst are host arrays, or device
 * arrays when on_device is nonzero (then no host copy is made).  A pair with
 * an id >= V or a self-loop fails with DCX_ERR_INVALID_ARG and names the
 * first such pair.  Replaces the host CSR build of Graph::readFile
 * (src/graph.cpp:146-250) followed by moveGraphToDevice
 * (src/cuda_memory.h:5-41). */
dcx_status dcx_graph_ingest_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                                  const uint32_t* dst, int32_t on_device, uint32_t slack,
                                  dcx_graph** out);
/* Ingest options (bit flags of dcx_graph_ingest). */
enum {
    DCX_INGEST_ON_DEVICE = 1,       /* src/dst are device arrays */
    DCX_INGEST_DROP_SELF_LOOPS = 2, /* skip pairs u->u, as Graph::readFile does
                                       (src/graph.cpp:167), instead of failing */
    DCX_INGEST_DEDUP = 4            /* keep each pair once; every list is then
                                       sorted by neighbour id */
};
/* dcx_graph_ingest_edges with options.  Without DCX_INGEST_DEDUP the graph is
 * the host build of the input (neighbours in input order, duplicates kept);
 * with it, the host build of the sorted, deduplicated pair list (what the
 * seeded generators return).  Ids >= V fail with DCX_ERR_INVALID_ARG, and so
 * do self-loops unless dropped.  The graph's E counts the kept pairs.
 * Replaces Graph::readFile's CSR build (src/graph.cpp:146-250) on raw pairs,
 * including the dedup a generator does before writing its file. */
dcx_status dcx_graph_ingest(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                            const uint32_t* dst, uint32_t flags, uint32_t slack,
                            dcx_graph** out);
/* The raw samples of dcx_gen_rmat (before its sort, self-loop drop and
 * dedup) drawn on the device into caller-allocated device arrays of
 * `samples` entries each; with dcx_graph_ingest(DCX_INGEST_ON_DEVICE |
 * DCX_INGEST_DROP_SELF_LOOPS | DCX_INGEST_DEDUP) they give the graph of
 * dcx_gen_rmat followed by dcx_graph_upload_edges without a host round trip.
 * No reference counterpart (the reference only loads files). */
dcx_status dcx_gen_rmat_device(dcx_ctx* ctx, uint32_t scale, uint64_t samples, double a,
                               double b, double c, uint64_t seed, uint32_t* src, uint32_t* dst);
dcx_status dcx_graph_free(dcx_graph* g);
/* V, current E, and kmax/lmax of the resident decomposition (-1 if none). */
dcx_status dcx_graph_info(const dcx_graph* g, uint32_t* V, uint64_t* E, int32_t* kmax,
                          int32_t* lmax);
/* Log of the passes of the resident decomposition, as the device recorded
 * them (no counterpart in the reference, which prints only the kmax/lmax
 * times, main.cu:256).  Entry 0 is the kmax pass (k = -1), entry 1+k row k;
 * rows not yet computed have variant 0.  These are the per-pass features
 * and timings the variant scheduler learns from. */
typedef struct dcx_pass_log {
    int32_t k;         /* -1 for the kmax pass */
    int32_t variant;   /* dcx_variant that ran the pass (0 = not run) */
    uint32_t members;  /* vertices of the pass (V, or V_k of the (k,0)-core) */
    uint32_t phases;   /* synchronised phases (peel), sweeps (h-index) or
                          refinement rounds (derived) */
    uint32_t levels;   /* level iterations (peel); 0 for h-index; derived:
                          1 for the first row of a chain (its time includes
                          the supports' set-up from the anchor row), 2 for
                          an anchor derived from row 0 (set-up included),
                          else 0 */
    uint32_t team;     /* workgroups that ran it */
    uint64_t edges;    /* edges of the pass (E, or E_k) */
    double ms;         /* device wall-clock time of the pass */
    uint32_t failed;   /* 1: stopped on inconsistent counters */
    uint32_t removed;  /* vertices removed when it ended */
    uint32_t alive;    /* alive-list length when it ended */
    int32_t level;     /* level when it ended */
} dcx_pass_log;
/* One synchronised phase of the last peel pass run on the whole grid (the
 * kmax pass, or a row of the plain variant) when cfg.trace_phases is set:
 * the per-iteration frontier statistics the scheduler learns from. */
typedef struct dcx_phase_trace {
    int32_t level;     /* level L of the phase */
    int32_t seed;      /* 1: seeding phase (alive-list scan), 0: cascade
                          (push), 2: cascade run as a pull */
    uint32_t removed;  /* vertices the phase removed */
    uint32_t alive;    /* alive-list length after it */
    double us;         /* device time since the previous phase */
} dcx_phase_trace;
dcx_status dcx_last_trace(dcx_ctx* ctx, dcx_phase_trace* out, uint32_t cap, uint32_t* n);
/* *n = kmax + 2 entries available; copies min(cap, *n) into out. */
dcx_status dcx_graph_pass_log(dcx_ctx* ctx, const dcx_graph* g, dcx_pass_log* out,
                              uint32_t cap, uint32_t* n);
/* Capacity of each adjacency array (E + free slots). */
dcx_status dcx_graph_capacity(const dcx_graph* g, uint64_t* out_cap, uint64_t* in_cap);
/* Copies the device CSR back (arrays sized V+1, V, capacity). */
dcx_status dcx_graph_download_csr(dcx_ctx* ctx, const dcx_graph* g, uint32_t* out_off,
                                  int32_t* out_deg, uint32_t* out_nbr, uint32_t* in_off,
                                  int32_t* in_deg, uint32_t* in_nbr);

/* In-core numbers (the kmax pass: KListInplace on the in/out-swapped graph
 * with k = 0, dcore main.cu:218-220; checkKmax main.cu:264-274).  kmaxes
 * (host, V entries) may be NULL; the values stay resident on the graph. */
dcx_status dcx_kmax(dcx_ctx* ctx, dcx_graph* g, int32_t* kmaxes, int32_t* kmax);

/* One l-row for a given k (KList / KListInplace, main.cu:144-195): lvals[v]
 * = max{l : v in D_{k,l}}, 0 outside the (k,0)-core; *lmax_k = last level.
 * variant: DCX_VARIANT_PEEL/AUTO (grid peel) or DCX_VARIANT_HINDEX. */
dcx_status dcx_klist(dcx_ctx* ctx, dcx_graph* g, int32_t k, int32_t variant, int32_t* lvals,
                     int32_t* lmax_k);

/* dcx_klist on the whole graph (KList semantics, main.cu:144-169) without
 * any device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  lvals is a caller-allocated
 * DEVICE array of V int32 receiving the row; *lmax_k the last level.  The
 * pass scratch is a static device array sized for at most 262,144 vertices
 * (12 MiB per device, held until the process ends); a larger graph gets
 * DCX_ERR_CAPACITY, and the phase trace (cfg.trace_phases) DCX_ERR_STATE.
 * All calls on one GPU share that array and the context's counter words:
 * calls that overlap on one GPU (another thread, another context) race. */
dcx_status dcx_klist_noalloc(dcx_ctx* ctx, dcx_graph* g, int32_t k, int32_t* lvals,
                             int32_t* lmax_k);

/* Full decomposition (dcore, main.cu:197-262; checkDCore main.cu:276-293):
 * kmaxes and rows k = 0..kmax stay resident on the device graph. */
dcx_status dcx_decompose(dcx_ctx* ctx, dcx_graph* g, int32_t variant, int32_t* kmax,
                         int32_t* lmax);

/* Result download: kmaxes[V]; dense rows k0..k0+nrows-1 (row-major, V
 * each, the .decomp row layout); ragged per-vertex layout (row_ptr[V+1],
 * values = lmaxes[0..kmaxes[v]][v], the .decompCPU order). */
dcx_status dcx_get_kmaxes(dcx_ctx* ctx, const dcx_graph* g, int32_t* kmaxes);
dcx_status dcx_get_rows(dcx_ctx* ctx, const dcx_graph* g, int32_t k0, int32_t nrows,
                        int32_t* rows);
/* ------------------------------------------------------------------ */
/* Batches of graphs.                                                 */
/* ------------------------------------------------------------------ */

/* Uploads n graphs (edge lists, ids local to each graph) as one device
 * graph, their disjoint union: graph i occupies ids [vbase[i], vbase[i+1])
 * (vbase has n+1 entries, filled here).  dcx_decompose on it decomposes
 * every graph in the same launches; graph i's kmaxes and rows are the
 * union's restricted to its range (rows above its own kmax are zero).
 * No reference counterpart: the reference decomposes one graph per run. */
dcx_status dcx_graph_upload_batch(dcx_ctx* ctx, uint32_t n, const uint32_t* V, const uint64_t* E,
                                  const uint32_t* const* src, const uint32_t* const* dst,
                                  uint32_t slack, uint32_t* vbase, dcx_graph** out);
/* Per-graph kmax and lmax of a decomposed batch (-1 / 0 for an empty graph). */
dcx_status dcx_batch_summary(dcx_ctx* ctx, const dcx_graph* g, uint32_t n, const uint32_t* vbase,
                             int32_t* kmax, int32_t* lmax);

/* ---- continuous-batching maintenance server ---------------------------
 * Replaces the reference's one-batch-at-a-time maintenance loop
 * (maintenanceFast, main.cu:1665-1751, driven from main.cu:1905-1924) with
 * a server that owns one decomposed graph and applies a stream of update
 * requests.  Admission: a request is staged in a fixed HBM pool
 * (hbm_budget) and uploaded on a copy stream as soon as it fits and an
 * in-flight slot is free, overlapping the maintenance running on the
 * device.  The worker applies every admitted request's edits in arrival
 * order (each checked against the graph as it then is, so each request's
 * status is what it would get alone) and then maintains kmaxes and rows
 * once for the group.  While served, the graph refuses other API calls
 * (DCX_ERR_STATE); results not polled before dcx_serve_stop are dropped. */
typedef struct dcx_server dcx_server;
typedef struct dcx_serve_config {
    uint64_t hbm_budget;       /* bytes of device staging for admitted requests (64 MiB) */
    uint32_t max_inflight;     /* admitted requests at most (4096) */
    uint32_t max_group;        /* requests maintained together at most (256; 1 = one by one) */
    uint32_t max_group_edges;  /* edits per group at most, unless one request has more (65536) */
    int32_t hash_states;       /* 1: each result carries dcx_state_hash after its group */
    uint64_t reserved[4];
} dcx_serve_config;
typedef struct dcx_serve_result {
    uint64_t id;               /* from dcx_serve_submit */
    int32_t status;            /* DCX_OK, or the error the request alone would get */
    uint32_t group;            /* maintenance group that applied it (0: rejected at submit) */
    uint32_t inserted, deleted;
    int32_t kmax_after;        /* kmax after its group */
    int32_t rows_recomputed;   /* rows its group recomputed */
    double submit_ms, admit_ms, start_ms, done_ms;  /* server clock (ms since start) */
    uint64_t state_hash;       /* after its group, if hash_states */
} dcx_serve_result;
typedef struct dcx_serve_stats {
    uint64_t requests, groups;
    uint32_t max_group, max_waiting;
    double busy_ms, elapsed_ms;
    uint64_t reserved[4];
} dcx_serve_stats;
dcx_status dcx_serve_config_init(dcx_serve_config* cfg);
/* Starts a server on a decomposed graph (NULL cfg = defaults). */
dcx_status dcx_serve_start(dcx_ctx* ctx, dcx_graph* g, const dcx_serve_config* cfg,
                           dcx_server** out);
/* Queues one update request (deletions then insertions, the dcx_update
 * contract); never waits for the device.  DCX_ERR_BUDGET: the request
 * alone exceeds hbm_budget. */
dcx_status dcx_serve_submit(dcx_server* sv, const uint32_t* ins, uint32_t n_ins,
                            const uint32_t* del, uint32_t n_del, uint64_t* id);
/* Moves up to cap finished results to out (waiting up to wait_ms for one). */
dcx_status dcx_serve_poll(dcx_server* sv, dcx_serve_result* out, uint32_t cap, uint32_t* n,
                          int32_t wait_ms);
/* Detail of the server's first failure ("" if none).  After a failure the
 * server applies nothing more: later requests complete with that status. */
const char* dcx_serve_last_error(dcx_server* sv);
/* Drains every submitted request, stops and frees the server. */
dcx_status dcx_serve_stop(dcx_server* sv, dcx_serve_stats* stats);
/* Hash of the resident decomposition: the 64-bit sum over v and k = -1..kmax
 * of mix64(v * 0x9E3779B97F4A7C15 + (k + 1) * 0xC2B2AE3D27D4EB4F + value),
 * value = kmaxes[v] for k = -1 and row k otherwise (mix64 = the SplitMix64
 * finaliser).  Independent of evaluation order; no reference counterpart
 * (the reference compares whole arrays on the host, src/utils.h:13-43). */
dcx_status dcx_state_hash(dcx_ctx* ctx, dcx_graph* g, uint64_t* hash);
/* Device-side compare, the verified counterpart of the reference's
 * compareArray kernel (src/cuda_utils.cuh:79-93, which counts mismatching
 * entries of two device arrays): a and b are device pointers of n int32;
 * *mismatches = entries that differ, *first (nullable) = the first such
 * index (UINT64_MAX if none). */
dcx_status dcx_compare_i32(dcx_ctx* ctx, const int32_t* a, const int32_t* b, uint64_t n,
                           uint64_t* mismatches, uint64_t* first);
/* Compares two resident decompositions on the device -- the reference's
 * "maintained == recomputed" check (main.cu:1905-1924, 2194-2211) without a
 * download: cell index v for kmaxes[v], V * (1 + k) + v for row k (a row
 * only one side has counts its nonzero cells).  Both graphs need the same
 * V and valid rows. */
dcx_status dcx_compare_decompositions(dcx_ctx* ctx, dcx_graph* a, dcx_graph* b,
                                      uint64_t* mismatches, uint64_t* first);

/* ---- learned variant scheduler ---------------------------------------
 * A small fp32 MLP (8 inputs, two hidden layers of 32 ReLU units, one
 * output) predicts log2(ms) of a team row of the batched launch from its
 * features: sizes and density of the (k,0)-core, the vertices that left
 * the core since row k-1, the kind of the row (peeled; derived from the
 * row above as the first row of a chain, whose time includes the
 * supports' set-up; derived later in a chain; derived from row 0 as an
 * anchor), the phase or round count and list reads of its last run of
 * that kind, and the team width it would run on.  Forward, backward and
 * Adam run in hand-written HIP kernels.  Attached to a context (and
 * cfg.chain_rows = 0), it chooses per team row among the exact variants --
 * peel it, derive it from row 0, or derive it from the row above by
 * h-index refinement -- and the team widths: the plan whose predicted
 * makespan is smallest, each prediction scaled by the ratio of measured to
 * predicted time of the row's last run of that kind on this graph.
 * Without it every chain_rows-th row from row 0 is an anchor of the kind
 * cfg.anchors names and the widths follow each unit's share of the work.
 * Every plan gives the same bits.  No reference counterpart (the
 * reference's schedule is fixed, main.cu:197-262). */
typedef struct dcx_mlp dcx_mlp;
uint32_t dcx_mlp_param_count(void);
/* Parameter layout: W1[32][8], b1[32], W2[32][32], b2[32], W3[32], b3. */
dcx_status dcx_mlp_create(dcx_ctx* ctx, uint64_t seed, dcx_mlp** out);
dcx_status dcx_mlp_free(dcx_mlp* m);
dcx_status dcx_mlp_get_params(dcx_mlp* m, float* out);
dcx_status dcx_mlp_set_params(dcx_mlp* m, const float* in);
/* y[s] = f(X[s*8 .. s*8+8)) on the device. */
dcx_status dcx_mlp_predict(dcx_mlp* m, const float* X, uint32_t n, float* y);
/* grad = sum_s d/dp 0.5 (f(X_s) - T_s)^2, *loss = sum_s 0.5 (f(X_s) - T_s)^2. */
dcx_status dcx_mlp_gradient(dcx_mlp* m, const float* X, const float* T, uint32_t n, float* grad,
                            float* loss);
/* `epochs` full-batch Adam steps (beta 0.9 / 0.999) on the mean of the above. */
dcx_status dcx_mlp_train(dcx_mlp* m, const float* X, const float* T, uint32_t n, uint32_t epochs,
                         float lr, float* final_loss);
/* Attach (or with NULL detach) the model as the batched rows scheduler.
 * Per graph, the fastest plan measured since the graph's last edit is kept
 * and reused unless the model predicts a new plan at least 3 % faster. */
dcx_status dcx_set_scheduler(dcx_ctx* ctx, dcx_mlp* m);
/* Training samples from the resident decomposition's pass log (the team
 * rows of the batched launch, peeled or derived): features (with_log = 0:
 * phase/round count set to unknown) and targets log2(ms + 0.01); *n =
 * samples available. */
dcx_status dcx_graph_pass_features(dcx_ctx* ctx, const dcx_graph* g, int32_t with_log,
                                   uint32_t cap, float* X, float* T, uint32_t* n);

/* Autotuner for the batched rows launch (no reference counterpart: the
 * reference's launch shape is fixed, src/common.h:38-56).  Times the rows
 * launch on `g` for candidate shapes (workgroups, team_max, shrink_at) by
 * coordinate descent, best of `reps` runs each (0 = 2), and leaves the
 * fastest configured in ctx with g's rows computed by it.  Every shape
 * gives identical results; only the time differs. */
typedef struct dcx_tune_result {
    int32_t rows_slots;   /* chosen workgroups of the launch */
    int32_t team_max;     /* chosen team cap */
    int32_t shrink_at;    /* chosen narrowing threshold */
    int32_t candidates;   /* shapes timed */
    double best_ms;       /* rows kernel time of the chosen shape */
    double default_ms;    /* rows kernel time of the default shape */
    int32_t lds_members;  /* chosen LDS-resident row cap (-1: no LDS rows), as
                             dcx_config.lds_members takes it */
    int32_t reserved32;
    double reserved[3];
} dcx_tune_result;
dcx_status dcx_autotune(dcx_ctx* ctx, dcx_graph* g, int32_t reps, dcx_tune_result* out);

/* ------------------------------------------------------------------ */
/* Edge updates and maintenance.                                      */
/* ------------------------------------------------------------------ */

/* Per-batch report.  The insert/kmax/rows split mirrors the
 * "insert: kmax: klist:" timings of maintenanceFast (main.cu:1665-1751). */
typedef struct dcx_update_stats {
    uint32_t inserted, deleted;
    uint32_t relocated;      /* vertex lists moved to make room */
    int32_t m_del;           /* max min(kmax u, kmax v) over deleted edges, old kmaxes (-1: none) */
    int32_t m_ins;           /* same over inserted edges, new kmaxes (-1: none) */
    int32_t rows_recomputed; /* l-rows k = 0..rows_recomputed-1 were recomputed */
    int32_t kmax_before, kmax_after;
    double insert_ms, kmax_ms, rows_ms, total_ms;
    int32_t path;            /* 0: recomputation (maintenance = 1, or the budgeted mode
                                backing off after a batch given up), 1: local
                                maintenance, 2: local maintenance given up for
                                recomputation (budget exceeded or worklist overflow) */
    int32_t net_deleted, net_inserted; /* the batch's net effect (an edge deleted and
                                inserted again in one batch changes nothing) */
    int32_t reserved[5];
} dcx_update_stats;

/* Applies a batch -- deletions first, then insertions; pairs are (u,v)
 * interleaved like the reference's modified_edges (main.cu:351-352) -- and
 * maintains the resident kmaxes and l-rows so that they equal a full
 * recomputation.  Replaces maintenance / maintenanceFast (main.cu:1536-1751)
 * with kmaxMaintenance(Delete) (main.cu:668-855).  Errors leave the graph
 * unchanged: an id out of range, a self-loop or a pair listed twice
 * (DCX_ERR_INVALID_ARG), an inserted edge already present
 * (DCX_ERR_INVALID_ARG), a deleted edge absent (DCX_ERR_NOT_FOUND). */
dcx_status dcx_update(dcx_ctx* ctx, dcx_graph* g, const uint32_t* ins, uint32_t n_ins,
                      const uint32_t* del, uint32_t n_del, dcx_update_stats* stats);

/* Work counters of the last local maintenance on a context (the figures the
 * reference prints under PRINT_MAINTENANCE_STATS, main.cu:1552-1564, for
 * this build's tasks): out[4 t + i] for task type t = 1 kmax deletion,
 * 2 row deletion, 3 kmax/row-0 insertion rounds, 4 row insertion; i = 0
 * vertices evaluated, 1 candidates taken in, 2 list entries the
 * evaluations covered, 3 list entries of expanded candidates.  n >= 20. */
dcx_status dcx_last_maintenance_counters(const dcx_ctx* ctx, uint64_t* out, uint32_t n);

/* Structural edits only (the resident decomposition is dropped):
 * GraphGPU::insertEdges (src/graph_gpu.cu:71-74, graphInsertEdges :49-69) /
 * Graph::insertEdgesInPlace (src/graph.cpp:63-70), and the delete methods
 * that the reference leaves unimplemented (src/graph.cpp:73-108,
 * src/graph_gpu.h:18-19).  Lists that run out of free slots are moved. */
dcx_status dcx_graph_insert_edges(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                  uint32_t n);
dcx_status dcx_graph_delete_edges(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                  uint32_t n);

/* Host-side batching of edges by root kmax (getKmaxEdgeBatches,
 * main.cu:1441-1525): mode 0 = KEDGE_BATCHING (distinct roots per batch),
 * mode 1 = KMAX_BATCHING.  batch_of_edge[n] receives each edge's batch. */
dcx_status dcx_kmax_edge_batches(uint32_t V, const int32_t* kmaxes, const uint32_t* pairs,
                                 uint32_t n, int32_t mode, uint32_t* batch_of_edge,
                                 uint32_t* n_batches);

/* The same batches computed on the device from the graph's resident kmaxes
 * (no kmax download): KEDGE by a stable sort of the edges by root and their
 * rank among the root's edges; KMAX by one workgroup playing the rounds.
 * pairs and batch_of_edge are host arrays as in dcx_kmax_edge_batches,
 * which this equals edge for edge.  Replaces getKmaxEdgeBatches
 * (main.cu:1441-1525) and its getKmaxFromDeviceMemory round trip
 * (main.cu:1713). */
dcx_status dcx_kmax_edge_batches_device(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                        uint32_t n, int32_t mode, uint32_t* batch_of_edge,
                                        uint32_t* n_batches);

/* Number of ragged values (sum over v of kmaxes[v]+1). */
dcx_status dcx_ragged_size(dcx_ctx* ctx, const dcx_graph* g, uint64_t* n);
dcx_status dcx_get_ragged(dcx_ctx* ctx, const dcx_graph* g, int64_t* row_ptr, int32_t* values);

#ifdef __cplusplus
}
#endif
#endif /* DCX_H */
