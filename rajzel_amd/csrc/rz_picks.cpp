/*
 * rz_picks.cpp -- pick files of calc_FA_res.c:28-37, parsed once per change
 * (rz_read_picks of rajzel_amd.h).
 *
 * Each file read is kept under its name in a small hash table: the stat()
 * identity it had (device, inode, size, modification time), the ntr it was
 * read for, and for every receiver the parsed value and whether fscanf
 * succeeded.  A hit copies the parsed values into Tobs and leaves the
 * others, which is what parsing the unchanged file again does.  Every
 * receiver gets its own fscanf call, as in the reference: after a failed
 * conversion a later one can still succeed (glibc consumes "1e+" before it
 * fails), so a parse does not stop at the first failure.
 */
#include "../../include/rajzel_amd.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

typedef struct rz_pick_entry {
    char *name;
    dev_t dev;
    ino_t ino;
    off_t size;
    struct timespec mtim;
    int ntr;
    float *val;             /* ntr values; those with ok[i] == 0 are unused */
    unsigned char *ok;      /* fscanf succeeded for receiver i + 1 */
    struct rz_pick_entry *next;
} rz_pick_entry;

static constexpr int kBuckets = 1024;
static rz_pick_entry *g_bucket[kBuckets];
static long g_parses = 0;
static int g_enabled = -1;  /* RAJZEL_AMD_PICK_CACHE, read on first use */

static unsigned bucket_of(const char *s)
{
    unsigned h = 2166136261u;   /* FNV-1a */
    for (; *s; ++s) h = (h ^ (unsigned char)*s) * 16777619u;
    return h % kBuckets;
}

static int same_file(const rz_pick_entry *e, const struct stat *st, int ntr)
{
    return e->ntr == ntr && e->dev == st->st_dev && e->ino == st->st_ino && e->size == st->st_size &&
           e->mtim.tv_sec == st->st_mtim.tv_sec && e->mtim.tv_nsec == st->st_mtim.tv_nsec;
}

int rz_read_picks(const char *name, float *Tobs, int ntr)
{
    if (g_enabled < 0) {
        const char *v = getenv("RAJZEL_AMD_PICK_CACHE");
        g_enabled = !(v && v[0] == '0');
    }
    struct stat st;
    const int have_st = stat(name, &st) == 0;
    const unsigned b = bucket_of(name);
    rz_pick_entry *e = g_bucket[b];
    while (e && strcmp(e->name, name) != 0) e = e->next;
    if (g_enabled && have_st && e && same_file(e, &st, ntr)) {
        for (int i = 0; i < ntr; ++i)
            if (e->ok[i]) Tobs[i + 1] = e->val[i];
        return 0;
    }
    FILE *fp = fopen(name, "r");
    if (fp == NULL) return -1;
    const size_t n = ntr > 0 ? (size_t)ntr : 1;
    float *val = (float *)malloc(n * sizeof(float));
    unsigned char *ok = (unsigned char *)malloc(n);
    for (int i = 1; i <= ntr; ++i) {
        const int got = fscanf(fp, "%e", &Tobs[i]) == 1;   /* Tobs[i] keeps its value otherwise */
        if (val && ok) {
            ok[i - 1] = (unsigned char)got;
            val[i - 1] = Tobs[i];
        }
    }
    fclose(fp);
    ++g_parses;
    if (!g_enabled || !have_st || !val || !ok) {   /* nothing to keep (or no memory for it) */
        free(val);
        free(ok);
        return 0;
    }
    if (!e) {
        e = (rz_pick_entry *)calloc(1, sizeof *e);
        char *nm = strdup(name);
        if (!e || !nm) {
            free(e); free(nm); free(val); free(ok);
            return 0;
        }
        e->name = nm;
        e->next = g_bucket[b];
        g_bucket[b] = e;
    } else {
        free(e->val);
        free(e->ok);
    }
    e->dev = st.st_dev; e->ino = st.st_ino; e->size = st.st_size; e->mtim = st.st_mtim;
    e->ntr = ntr; e->val = val; e->ok = ok;
    return 0;
}

long rz_picks_parses(void) { return g_parses; }

void rz_picks_clear(void)
{
    for (int b = 0; b < kBuckets; ++b) {
        rz_pick_entry *e = g_bucket[b];
        while (e) {
            rz_pick_entry *n = e->next;
            free(e->name); free(e->val); free(e->ok); free(e);
            e = n;
        }
        g_bucket[b] = NULL;
    }
}
