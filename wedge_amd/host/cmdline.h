/*
 * cmdline.h -- command-line options of the wedge front end.
 *
 * Keeps the reference's options and defaults (include/cmdline.h:26-52,
 * source/cmdline.c:65-108): -a app, -i graph, -N iterations, -n threads,
 * -u NUMA nodes, -r root, -t threshold, -o output, -h, -V.  -n and -u select
 * CPU threads and NUMA nodes in the reference; here they are accepted and
 * ignored (one GPU does the work).  Additions: -d device, -s schedule,
 * -b block granularity, -D damping, -f FP32 PageRank contributions,
 * -g generator, -W weight type of the input's third word, -m schedule model,
 * -R list of roots (BFS / SSSP from several sources at once) or seeds
 * (application ppr).  A comma-separated -i or -g names the members of a batch
 * of graphs run together.  -c validates the result on the device; -P writes
 * ranks at full precision.
 */
#ifndef WG_CMDLINE_H
#define WG_CMDLINE_H

#include <stdint.h>

#define CMDLINE_DEFAULT_NUM_ITERATIONS 1
#define CMDLINE_DEFAULT_ROOT_VERTEX 0
#define CMDLINE_DEFAULT_FRONTIER_THRESHOLD_PCT 20
#define CMDLINE_MAX_ROOTS 64

typedef struct cmdline_opts_t {
    char graph_input_filename[16384]; /* -i (comma-separated: a batch) */
    char* graph_application_output_filename; /* -o */
    char application[64];            /* -a */
    uint32_t num_iterations;         /* -N */
    uint32_t num_threads;            /* -n (accepted, unused) */
    uint64_t root_vertex;            /* -r */
    uint64_t frontier_threshold_pct; /* -t */
    int device;                      /* -d */
    int schedule;                    /* -s plain|learned|dense|frontier */
    int block_log2;                  /* -b */
    double damping;                  /* -D */
    int contrib_fp32;                /* -f */
    char generator[16384];           /* -g rmat:SCALE:EF:SEED[:sym] | grid:ROWS:COLS:SEED (comma-separated: a batch) */
    int weight_kind;                 /* -W int|f64 (SSSP input third word) */
    char model[1024];                /* -m */
    int root_given;
    uint64_t roots[64];              /* -R */
    uint32_t nroots;
    int validate;                    /* -c */
    int full_precision;              /* -P */
} cmdline_opts_t;

/* Parses argv; prints a message and exits with the reference's codes
 * (1 unknown option, 2 invalid value, 3 missing value, 4 extraneous value,
 * 5 missing required option, 6 incompatible options) on error. */
void cmdline_parse_options_or_die(int argc, char* argv[]);
const cmdline_opts_t* cmdline_get_current_settings(void);

#endif /* WG_CMDLINE_H */
