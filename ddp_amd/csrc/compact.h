// Device-side compaction of an update log into CSR rows (compact.hip).
#pragma once

#include "ddp_internal.h"

// A graph version to compact: base rows off[N+1] / col / w, where edge j is
// live iff col[j] >= 0 and (dver == null or dver[j] > ver), plus inserted
// edges on per-source lists (lhead[u], lnext[x]; links pack entry << 32 |
// inserting version, ~0 ends a list) whose entry x is live iff its link
// version <= ver < ldel[x].  lhead may be null (no inserted edges).
struct MergeSrc {
    const int32_t *off, *col;
    const uint32_t *w;
    int32_t N;
    const int32_t *dver;
    int32_t ver;
    const uint64_t *lhead, *lnext;
    const int32_t *lv;
    const uint32_t *lw;
    const int32_t *ldel;
};

// Device CSR owned by the caller after a successful call (ddp_merged_free).
struct MergedCsr {
    int32_t *off = nullptr;  // N + 1
    int32_t *col = nullptr;
    uint32_t *w = nullptr;
    int32_t *eid = nullptr;  // transposed rows: the forward edge id of each entry
    int32_t N = 0, E = 0;
};

// The merged rows of m at version m.ver: each row's live base edges in row
// order, then its live inserted edges oldest first.
int ddp_merge_rows(ddp_ctx *ctx, const MergeSrc &m, MergedCsr *out);
// The transposed rows of f (the order inside a row is not fixed), with
// forward edge ids when with_eid.
int ddp_transpose_rows(ddp_ctx *ctx, const MergedCsr &f, bool with_eid, MergedCsr *out);
void ddp_merged_free(MergedCsr *c);
// out[0..n] = exclusive prefix sums of cnt[0..n) (out[n] = total).
int ddp_scan_counts(ddp_ctx *ctx, const int32_t *cnt, int64_t n, int32_t *out, int64_t *total);
