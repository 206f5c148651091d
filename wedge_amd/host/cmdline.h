/*
 * cmdline.h -- command-line options of the wedge front end.
 *
 * Keeps the reference's options and defaults (include/cmdline.h:26-52,
 * source/cmdline.c:65-108): -a app, -i graph, -N iterations, -n threads,
 * -u NUMA nodes, -r root, -t threshold, -o output, -h, -V.  -n and -u select
 * CPU threads and NUMA nodes in the reference; here they are accepted and
 * ignored (one GPU does the work).  Additions: -d device, -s schedule,
 * -b block granularity, -D damping, -f FP32 PageRank contributions,
 * -g generator, -W weight type of the input's third word, -m schedule model.
 */
#ifndef WG_CMDLINE_H
#define WG_CMDLINE_H

#include <stdint.h>

#define CMDLINE_DEFAULT_NUM_ITERATIONS 1
#define CMDLINE_DEFAULT_ROOT_VERTEX 0
#define CMDLINE_DEFAULT_FRONTIER_THRESHOLD_PCT 20

typedef struct cmdline_opts_t {
    char graph_input_filename[1024]; /* -i */
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
    char generator[256];             /* -g rmat:SCALE:EF:SEED[:sym] | grid:ROWS:COLS:SEED */
    int weight_kind;                 /* -W int|f64 (SSSP input third word) */
    char model[1024];                /* -m */
    int root_given;
} cmdline_opts_t;

/* Parses argv; prints a message and exits with the reference's codes
 * (1 unknown option, 2 invalid value, 3 missing value, 4 extraneous value,
 * 5 missing required option, 6 incompatible options) on error. */
void cmdline_parse_options_or_die(int argc, char* argv[]);
const cmdline_opts_t* cmdline_get_current_settings(void);

#endif /* WG_CMDLINE_H */
