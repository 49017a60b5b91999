/*
 * swm_index -- index builder CLI with serial_index.c's options
 * (getopt "h::i:s::p:", serial_index.c:819-846): -i reference FASTA,
 * -p index prefix, -s seed length (attached, default 12), -h help.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "swm.h"

static void usage(void)
{
    printf("-i\tinput reference\n-p\tprefix index name\n-s\t[opt, default:12] seed length\n-h\tprint help information\n");
}

int main(int argc, char **argv)
{
    const char *ref = NULL, *prefix = NULL;
    long k = 12;
    int opt;
    while ((opt = getopt(argc, argv, "h::i:s::p:")) != -1) {
        switch (opt) {
        case 'i': ref = optarg; break;
        case 'p': prefix = optarg; break;
        case 's': if (optarg) k = strtol(optarg, NULL, 10); break;
        case 'h': usage(); return 0;
        default: usage(); return 2;
        }
    }
    if (!ref || !prefix || k < 2 || k > 14) { usage(); return 2; }
    int rc = swm_index_build(ref, prefix, (uint32_t)k);
    if (rc != SWM_OK) { fprintf(stderr, "swm_index: build failed (status %d)\n", rc); return 1; }
    return 0;
}
