/* A C client of the ABI in include/meikey_hip.h, as meikey's Rust FFI would use it:
 * open device 0, compile a pattern, run the Searcher until its first exact hit, and print
 * the index, public key and OpenSSH line.
 *   mk_example PATTERN [KEY_HEX64]    (case-insensitive suffix; a fixed key when given) */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "meikey_hip.h"

static int fail(const char* what) {
    char msg[512];
    mk_last_error(msg, sizeof msg);
    fprintf(stderr, "%s: %s\n", what, msg);
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s PATTERN [KEY_HEX64]\n", argv[0]);
        return 2;
    }
    uint8_t key[32];
    int fixed = 0;
    if (argc > 2 && strlen(argv[2]) == 64) {
        for (int i = 0; i < 32; i++) {
            unsigned v = 0;
            sscanf(argv[2] + 2 * i, "%2x", &v);
            key[i] = (uint8_t)v;
        }
        fixed = 1;
    }
    mk_options opt;
    mk_default_options(&opt);
    opt.width = 16; /* a 50 MB table: quick to build for an example */
    mk_gpu* g = NULL;
    if (mk_open(0, &opt, &g) != MK_OK) return fail("mk_open");
    mk_pattern* p = NULL;
    int kind = 0;
    if (mk_pattern_new(argv[1], (uint32_t)strlen(argv[1]), 0, MK_MODE_SUFFIX, &p, &kind) != MK_OK) {
        char msg[256];
        mk_pattern_last_error(msg, sizeof msg);
        fprintf(stderr, "pattern: %s\n", msg);
        mk_close(g);
        return 1;
    }
    mk_search_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.pattern = p;
    cfg.seed_source = MK_SEED_CHACHA20;
    cfg.key = fixed ? key : NULL;
    mk_outcome out;
    int rc = 0;
    if (mk_search(g, &cfg, NULL, NULL, &out) != MK_OK) {
        rc = fail("mk_search");
    } else if (!out.found) {
        fprintf(stderr, "no key found\n");
        rc = 1;
    } else {
        char line[256];
        size_t n = 0;
        if (mk_openssh_public(out.pk, "example", line, sizeof line, &n) != MK_OK) return fail("mk_openssh_public");
        printf("index %llu\npk ", (unsigned long long)out.index);
        for (int i = 0; i < 32; i++) printf("%02x", out.pk[i]);
        printf("\ntries %llu\n%s\n", (unsigned long long)out.stats.tries, line);
    }
    mk_pattern_free(p);
    mk_close(g);
    return rc;
}
