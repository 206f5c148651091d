/*
 * cmdline.c -- option parsing for the wedge front end.
 *
 * Same switch syntax and error behaviour as the reference (source/cmdline.c):
 * every option is '-' plus one letter; a following argument that does not
 * start with '-' is its value; each kind of error prints a message and exits
 * with the reference's code.
 */
#define _GNU_SOURCE
#include "cmdline.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wedge_host.h"

static cmdline_opts_t opts;

static int is_recognized(char c)
{
    return strchr("ahinNortuVdsbDfgWmRcP", c) != NULL;
}

static int requires_value(char c)
{
    return strchr("ainNortudsbDgWmR", c) != NULL;
}

static void die_unknown(const char* argv0, const char* opt)
{
    printf("%s: Unrecognized option `%s'.\n", argv0, opt);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(1);
}

static void die_invalid(const char* argv0, const char* opt, const char* val)
{
    printf("%s: Invalid value `%s' for option `%s'.\n", argv0, val, opt);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(2);
}

static void die_missing_value(const char* argv0, const char* opt)
{
    printf("%s: Missing argument for option `%s'.\n", argv0, opt);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(3);
}

static void die_extraneous(const char* argv0, const char* opt)
{
    printf("%s: Option `%s' does not accept an argument.\n", argv0, opt);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(4);
}

static void die_missing_option(const char* argv0, const char* opt)
{
    printf("%s: Missing required option `-%s'.\n", argv0, opt);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(5);
}

static void die_incompatible(const char* argv0)
{
    printf("%s: Specified option combination is incompatible.\n", argv0);
    printf("Try `%s -h' for more information.\n", argv0);
    exit(6);
}

static void usage(const char* argv0)
{
    printf("Usage: %s [options] -a application-name -i input-graph\n", argv0);
    printf("       %s [options] -a application-name -g generator\n", argv0);
    printf("       %s -h\n       %s -V\n\n", argv0, argv0);
    printf("Required:\n");
    printf("  -a application-name   bfs | cc | pr | sssp (and the reference's aliases)\n");
    printf("  -i input-graph        binary edge list: u64 V, u64 E, then (src, dst[, weight]) grouped by destination\n\n");
    printf("Options:\n");
    printf("  -h                    prints this information and exits\n");
    printf("  -N num-iterations     PageRank iterations (default %d)\n", CMDLINE_DEFAULT_NUM_ITERATIONS);
    printf("  -n num-threads        accepted for compatibility; the GPU does the work\n");
    printf("  -u node[,node...]     accepted for compatibility; the GPU does the work\n");
    printf("  -o output-file        writes \"<vertex> <value>\" per line\n");
    printf("  -r root-vertex        BFS/SSSP root (default %d)\n", CMDLINE_DEFAULT_ROOT_VERTEX);
    printf("  -t threshold          percentage of active edges above which the dense pull is used (default %d)\n",
           CMDLINE_DEFAULT_FRONTIER_THRESHOLD_PCT);
    printf("  -V                    prints version information and exits\n");
    printf("  -d device             HIP device (default 0)\n");
    printf("  -s schedule           plain | learned | dense | frontier (default plain)\n");
    printf("  -b block-log2         Wedge granularity, 2^b edges per bit, %d..%d (default %d)\n", WG_MIN_BLOCK_LOG2,
           WG_MAX_BLOCK_LOG2, WG_DEFAULT_BLOCK_LOG2);
    printf("  -D damping            PageRank damping (default 0.875)\n");
    printf("  -f                    PageRank: FP32 contributions, FP64 accumulation\n");
    printf("  -g generator          rmat:SCALE:EF:SEED[:sym][:w] | grid:ROWS:COLS:SEED instead of -i\n");
    printf("  -W int|f64            SSSP input: third word is an integer (default) or an FP64 weight\n");
    printf("  -m model-file         schedule model for -s learned\n");
    printf("  -R v[,v...]           BFS/SSSP from up to %d roots in one traversal; seeds of -a ppr (up to 32)\n",
           CMDLINE_MAX_ROOTS);
    printf("  -a ppr                personalised PageRank from the -R seeds (-N iterations, -D damping)\n");
    printf("  -c                    validate the BFS/CC/SSSP result on the device (exit code 7 if invalid)\n");
    printf("  -P                    -o writes ranks with 17 significant digits instead of 5\n");
    printf("  -i a,b,... / -g a,b,...  a batch of graphs run together (BFS/SSSP from vertex -r of each)\n");
    exit(0);
}

static uint64_t parse_u64(const char* argv0, const char* opt, const char* val)
{
    char* end = NULL;
    const unsigned long long x = strtoull(val, &end, 10);
    if (!*val || *end != '\0')
        die_invalid(argv0, opt, val);
    return (uint64_t)x;
}

static void parse_one(const char* argv0, const char* opt, const char* val)
{
    if (opt[0] != '-' || opt[1] == '\0' || opt[2] != '\0' || !is_recognized(opt[1]))
        die_unknown(argv0, opt);
    if (requires_value(opt[1]) && !val)
        die_missing_value(argv0, opt);
    if (!requires_value(opt[1]) && val)
        die_extraneous(argv0, opt);
    switch (opt[1]) {
    case 'h':
        usage(argv0);
        break;
    case 'V':
        printf("wedge-mi355x v1.0 for Linux / gfx950\n");
        exit(0);
    case 'a':
        strncpy(opts.application, val, sizeof(opts.application) - 1);
        break;
    case 'i':
        strncpy(opts.graph_input_filename, val, sizeof(opts.graph_input_filename) - 1);
        break;
    case 'n':
        opts.num_threads = (uint32_t)parse_u64(argv0, opt, val);
        break;
    case 'u':
        for (const char* p = val; *p; ++p)
            if (!isdigit((unsigned char)*p) && *p != ',')
                die_invalid(argv0, opt, val);
        break;
    case 'N': {
        const uint64_t n = parse_u64(argv0, opt, val);
        if (n < 1 || n > 0xffffffffull)
            die_invalid(argv0, opt, val);
        opts.num_iterations = (uint32_t)n;
        break;
    }
    case 'o':
        opts.graph_application_output_filename = (char*)val;
        break;
    case 'r':
        opts.root_vertex = parse_u64(argv0, opt, val);
        opts.root_given = 1;
        break;
    case 't':
        opts.frontier_threshold_pct = parse_u64(argv0, opt, val);
        break;
    case 'd':
        opts.device = (int)parse_u64(argv0, opt, val);
        break;
    case 's':
        if (!strcasecmp(val, "plain"))
            opts.schedule = WG_SCHED_PLAIN;
        else if (!strcasecmp(val, "learned"))
            opts.schedule = WG_SCHED_LEARNED;
        else if (!strcasecmp(val, "dense"))
            opts.schedule = WG_SCHED_DENSE;
        else if (!strcasecmp(val, "frontier"))
            opts.schedule = WG_SCHED_FRONTIER;
        else
            die_invalid(argv0, opt, val);
        break;
    case 'b': {
        const uint64_t b = parse_u64(argv0, opt, val);
        if (b < WG_MIN_BLOCK_LOG2 || b > WG_MAX_BLOCK_LOG2)
            die_invalid(argv0, opt, val);
        opts.block_log2 = (int)b;
        break;
    }
    case 'D': {
        char* end = NULL;
        const double d = strtod(val, &end);
        if (*end != '\0' || !(d >= 0.0 && d <= 1.0))
            die_invalid(argv0, opt, val);
        opts.damping = d;
        break;
    }
    case 'f':
        opts.contrib_fp32 = 1;
        break;
    case 'c':
        opts.validate = 1;
        break;
    case 'P':
        opts.full_precision = 1;
        break;
    case 'g':
        strncpy(opts.generator, val, sizeof(opts.generator) - 1);
        break;
    case 'W':
        if (!strcasecmp(val, "int"))
            opts.weight_kind = WG_WEIGHT_U32;
        else if (!strcasecmp(val, "f64"))
            opts.weight_kind = WG_WEIGHT_F64;
        else
            die_invalid(argv0, opt, val);
        break;
    case 'm':
        strncpy(opts.model, val, sizeof(opts.model) - 1);
        break;
    case 'R': {
        opts.nroots = 0;
        const char* p = val;
        while (*p) {
            char* end = NULL;
            if (!isdigit((unsigned char)*p) || opts.nroots == CMDLINE_MAX_ROOTS)
                die_invalid(argv0, opt, val);
            opts.roots[opts.nroots++] = (uint64_t)strtoull(p, &end, 10);
            if (*end == ',' && end[1] != '\0')
                p = end + 1;
            else if (*end == '\0')
                p = end;
            else
                die_invalid(argv0, opt, val);
        }
        if (opts.nroots == 0)
            die_invalid(argv0, opt, val);
        break;
    }
    default:
        die_unknown(argv0, opt);
    }
}

void cmdline_parse_options_or_die(int argc, char* argv[])
{
    memset(&opts, 0, sizeof(opts));
    opts.num_iterations = CMDLINE_DEFAULT_NUM_ITERATIONS;
    opts.root_vertex = CMDLINE_DEFAULT_ROOT_VERTEX;
    opts.frontier_threshold_pct = CMDLINE_DEFAULT_FRONTIER_THRESHOLD_PCT;
    opts.schedule = WG_SCHED_PLAIN;
    opts.block_log2 = WG_DEFAULT_BLOCK_LOG2;
    opts.damping = 0.875;
    opts.weight_kind = WG_WEIGHT_U32;
    int i = 1;
    while (i < argc) {
        const char* opt = argv[i];
        const char* val = NULL;
        if (i + 1 < argc && isprint((unsigned char)argv[i + 1][0]) && argv[i + 1][0] != '-') {
            val = argv[i + 1];
            i += 1;
        }
        parse_one(argv[0], opt, val);
        i += 1;
    }
    if (opts.application[0] == '\0')
        die_missing_option(argv[0], "a");
    if (opts.graph_input_filename[0] == '\0' && opts.generator[0] == '\0')
        die_missing_option(argv[0], "i");
    if (opts.graph_input_filename[0] != '\0' && opts.generator[0] != '\0')
        die_incompatible(argv[0]);
}

const cmdline_opts_t* cmdline_get_current_settings(void)
{
    return &opts;
}
