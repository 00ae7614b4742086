/*
 * orc_receivers.c -- CPU oracle of receiver-driven probe selection
 * (TEST INFRASTRUCTURE; the checker, never the product).
 *
 * Restates the sparse selection of the reference's Split path:
 *   candidate permutation     permuteCandidateDomain / inverse / transportCandidateRank /
 *                             transportCandidateFromRank / transportCandidateSelected
 *                             (shaders/split.slang:567-663; the reference's own restatement
 *                             src/transport_tests.rs:1127-1231, src/render.rs:12954-13001)
 *   seed                      shaders/split.slang:734-793 (primary receivers; no secondary
 *                             seeds)
 *   rank                      shaders/split.slang:799-893, with its cold-priority phases
 *                             (coldPriorityEnabled / reserveColdUpdate, :685-705)
 *   parents / prefix          shaders/split.slang:898-961
 *   generateDirections        shaders/split.slang:963-983
 *
 * The reference resolves three orders with atomics (which descriptor of a
 * key becomes canonical, the order of the local ordinals, the order of the
 * roots on the root counter and of the trace tickets).  This restatement
 * fixes each to the order a sequential pass over ascending descriptor
 * index gives -- one of the outcomes the reference allows -- so the GPU can
 * be compared bit for bit: canonical = lowest descriptor index of the key,
 * ordinals in ascending descriptor index, roots in ascending canonical index,
 * tickets in ascending receiver index.
 *
 * Pinned by: the reference's candidate-schedule tests (bijection, exact
 * budget, the uneven-tail cursor values, inverse round trip, spatial
 * uniformity) and the probe-prefix fixture of
 * src/transport_tests.rs:1887-2480 (tests/test_oracle_receivers.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lux_oracle.h"

#define EMPTY 0xffffffffu
#define PROBE_LOD_MASK (31u << 18)
#define MAX_PROBE_LOD 24u

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline int finitef_(float f) { return (fbits(f) & 0x7f800000u) != 0x7f800000u; }

static inline uint32_t transport_hash(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

/* ---------------- candidate permutation ---------------- */
static uint32_t permute(uint32_t v, uint32_t mask, uint32_t seed) {
    v = (v ^ seed) & mask;
    v = (v ^ (v >> 16)) & mask;
    v = (v * 0x7feb352du) & mask;
    v = (v ^ (v >> 15)) & mask;
    v = (v * 0x846ca68bu) & mask;
    v = (v ^ (v >> 16)) & mask;
    return v;
}

static uint32_t inverse_xorshift(uint32_t v, uint32_t shift) {
    uint32_t r = v;
    for (uint32_t off = shift; off < 32u; off += shift) r ^= v >> off;
    return r;
}

static uint32_t inverse_permute(uint32_t v, uint32_t mask, uint32_t seed) {
    v = inverse_xorshift(v, 16u) & mask;
    v = (v * 0x43021123u) & mask;
    v = inverse_xorshift(v, 15u) & mask;
    v = (v * 0x1d69e2a5u) & mask;
    v = inverse_xorshift(v, 16u) & mask;
    return (v ^ seed) & mask;
}

static uint32_t perm_mask(uint32_t count) {
    uint32_t m = 0;
    while (m < count - 1u) m = (m << 1) | 1u;
    return m;
}

static uint32_t perm_seed(uint32_t w, uint32_t h, uint32_t count, uint32_t mask) {
    return ((w * 0x9e3779b9u) ^ (h * 0x85ebca6bu) ^ count) & mask;
}

uint32_t orc_candidate_rank(uint32_t sample, uint32_t width, uint32_t height, uint32_t count) {
    if (count <= 1u) return 0u;
    uint32_t mask = perm_mask(count), seed = perm_seed(width, height, count, mask);
    uint32_t r = permute(sample, mask, seed);
    while (r >= count) r = permute(r, mask, seed);
    return r;
}

uint32_t orc_candidate_from_rank(uint32_t rank, uint32_t width, uint32_t height, uint32_t count) {
    if (count <= 1u) return 0u;
    uint32_t mask = perm_mask(count), seed = perm_seed(width, height, count, mask);
    uint32_t s = inverse_permute(rank, mask, seed);
    while (s >= count) s = inverse_permute(s, mask, seed);
    return s;
}

/* coldCandidateRank / coldCandidateFromRank (split.slang:631-657): the same
 * permutation with a frame-seeded seed. */
static uint32_t cold_seed(uint32_t w, uint32_t h, uint32_t count, uint32_t frame, uint32_t mask) {
    return (perm_seed(w, h, count, mask) ^ (frame * 0x9e3779b9u)) & mask;
}

uint32_t orc_cold_candidate_rank(uint32_t sample, uint32_t width, uint32_t height, uint32_t count, uint32_t frame) {
    if (count <= 1u) return 0u;
    uint32_t mask = perm_mask(count), seed = cold_seed(width, height, count, frame, mask);
    uint32_t r = permute(sample, mask, seed);
    while (r >= count) r = permute(r, mask, seed);
    return r;
}

uint32_t orc_cold_candidate_from_rank(uint32_t rank, uint32_t width, uint32_t height, uint32_t count,
                                      uint32_t frame) {
    if (count <= 1u) return 0u;
    uint32_t mask = perm_mask(count), seed = cold_seed(width, height, count, frame, mask);
    uint32_t s = inverse_permute(rank, mask, seed);
    while (s >= count) s = inverse_permute(s, mask, seed);
    return s;
}

/* coldPopulationBin (split.slang:680-687). */
static uint32_t cold_bin(uint32_t population) {
    uint32_t bin = 0;
    while (population > 1u && bin < 31u) {
        population >>= 1;
        ++bin;
    }
    return bin;
}

int orc_candidate_selected(uint32_t sample, const orc_receiver_params* rp) {
    uint32_t count = rp->width * rp->height;
    if (rp->flags & ORC_RX_DISABLED) return 0;
    if (count == 0u || rp->budget == 0u || rp->budget >= count) return 1;
    uint32_t start = (rp->fair_cursor & 0x03ffffffu) % count;
    uint32_t quota = rp->fair_quota;
    uint32_t rank = orc_candidate_rank(sample, rp->width, rp->height, count);
    if (quota == 0u) return 0;
    return rank >= start ? rank - start < quota : count - start + rank < quota;
}

/* ---------------- seed ---------------- */
/* Eligible receiver: finite position and normal, |n|^2 > 0.01; unit normal. */
int orc_rx_surface(const float* r, float pos[3], float n[3]) {
    float x = r[3], y = r[4], z = r[5];
    float nn = (x * x + y * y) + z * z;
    if (!(finitef_(r[0]) && finitef_(r[1]) && finitef_(r[2]))) return 0;
    if (!(finitef_(x) && finitef_(y) && finitef_(z)) || !(nn > 0.01f)) return 0;
    float nl = sqrtf(nn);
    pos[0] = r[0]; pos[1] = r[1]; pos[2] = r[2];
    n[0] = x / nl; n[1] = y / nl; n[2] = z / nl;
    return 1;
}

/* cacheKey (common.slang:418-423) with direction 0. */
void orc_rx_cell(const float pos[3], float base_spacing, uint32_t level, uint32_t lod, int32_t key[4]) {
    uint32_t l = lod < MAX_PROBE_LOD ? lod : MAX_PROBE_LOD;
    float s = base_spacing * ldexpf(1.0f, (int)(level + l));
    for (int a = 0; a < 3; ++a) key[a] = (int32_t)floorf(pos[a] / s);
    key[3] = (int32_t)((level << 12) | (l << 18));
}

void orc_rx_seed(const float* receivers, const orc_receiver_params* rp, const orc_lod_view* view,
                 float base_spacing, uint32_t levels, int cold, int32_t* keys, uint32_t* flags, uint32_t* lods) {
    const uint32_t count = rp->width * rp->height;
    for (uint32_t s = 0; s < count; ++s) {
        float pos[3], n[3];
        int ok = orc_rx_surface(receivers + 6ull * s, pos, n);
        uint32_t low = 0, high = 0;
        float blend = 0.0f;
        if (ok && view) orc_lod_select(view, pos, &low, &high, &blend);
        uint32_t selected_lod = 0;
        if (high != low) {
            uint32_t bits = transport_hash(s ^ (rp->frame * 0x9e3779b9u));
            float phase = (float)(bits >> 8) * (1.0f / 16777216.0f);
            selected_lod = phase < blend ? 1u : 0u;
        }
        /* cold priority: every receiver's selected LOD is eligible (eligibleRay) */
        int ray = ok && (cold || orc_candidate_selected(s, rp));
        for (uint32_t v = 0; v < 2; ++v) {
            lods[v * count + s] = v ? high : low;
            for (uint32_t L = 0; L < levels; ++L) {
                uint64_t d = ((uint64_t)v * count + s) * levels + L;
                int32_t* k = keys + 4 * d;
                if (!ok || (v == 1 && high == low)) {
                    k[0] = k[1] = k[2] = k[3] = 0;
                    flags[d] = 0;
                    continue;
                }
                orc_rx_cell(pos, base_spacing, L, v ? high : low, k);
                flags[d] = ORC_RX_VALID | ((L == 0 && ray && v == selected_lod) ? ORC_RX_RAY : 0u);
            }
        }
    }
}

/* ---------------- rank, parents, prefix, generateDirections ---------------- */
static uint32_t directory_bucket(const int32_t k[4], uint32_t cap) {
    uint32_t w = (uint32_t)k[3] & (PROBE_LOD_MASK | (3u << 12));
    uint32_t h = (uint32_t)k[0] * 0x8da6b343u ^ (uint32_t)k[1] * 0xd8163841u ^ (uint32_t)k[2] * 0xcb1ab31fu ^
                 w * 0x165667b1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h & (cap - 1u);
}

static int same_probe(const int32_t* a, const int32_t* b) {
    const uint32_t m = PROBE_LOD_MASK | (0xfu << 12);
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && (((uint32_t)a[3] ^ (uint32_t)b[3]) & m) == 0u;
}

typedef struct {
    uint32_t cap, levels;
    uint32_t* slot; /* [levels][cap] owning descriptor index */
    const int32_t* keys;
} directory;

static uint32_t dir_find(const directory* D, const int32_t* key, uint32_t level) {
    uint32_t b = directory_bucket(key, D->cap);
    for (uint32_t a = 0; a < D->cap; ++a) {
        uint32_t idx = D->slot[(uint64_t)level * D->cap + b];
        if (idx == EMPTY) return EMPTY;
        if (same_probe(D->keys + 4ull * idx, key)) return idx;
        b = (b + 1u) & (D->cap - 1u);
    }
    return EMPTY;
}

static uint32_t dir_insert(directory* D, const int32_t* key, uint32_t level, uint32_t idx) {
    uint32_t b = directory_bucket(key, D->cap);
    for (uint32_t a = 0; a < D->cap; ++a) {
        uint32_t* s = &D->slot[(uint64_t)level * D->cap + b];
        if (*s == EMPTY) {
            *s = idx;
            return idx;
        }
        if (same_probe(D->keys + 4ull * *s, key)) return *s;
        b = (b + 1u) & (D->cap - 1u);
    }
    return EMPTY;
}

uint32_t orc_rx_directory_capacity(uint32_t count) {
    uint32_t v = (count * 4u > 1u ? count * 4u : 1u) - 1u;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1u;
}

/* The cold-priority phases of rank (split.slang:836-888) after the
 * descriptors are ranked: one nominee per probe (lowest frame-seeded rank
 * among its eligible descriptors), probes admitted by population bin within
 * the budget, then fair-window receivers fill what the budget leaves.  The
 * orders the reference leaves to its atomics are fixed: admission tickets
 * inside a bin and the fair fill go in the order of its rank-mapped dispatch
 * (ascending candidate rank, variant 0 first); local ordinals are 0 for the
 * nominee, then the fill receivers in ascending descriptor index.  mark: 2
 * admitted nominee, 1 other traced; ordinals and level-0 counts are
 * written. */
static void cold_phases(uint32_t count, uint32_t levels, const int32_t* keys, const uint32_t* flags,
                        const uint32_t* canonical, const orc_receiver_params* rp, uint8_t* mark, uint32_t* cnt,
                        uint32_t* ordinal) {
    const uint64_t nd = 2ull * count * levels;
    const uint32_t budget = rp->budget;
    uint32_t* prio = (uint32_t*)calloc((size_t)nd, sizeof(uint32_t));
    uint32_t* pop = (uint32_t*)calloc((size_t)nd, sizeof(uint32_t));
    /* phase 0: nominee priority and population per level-0 probe */
    for (uint64_t d = 0; d < nd; d += levels) {
        if (!mark[d]) continue;
        uint32_t P = canonical[d], smp = (uint32_t)((d / levels) % count);
        uint32_t pr = 0xffffffffu - orc_cold_candidate_rank(smp, rp->width, rp->height, count, rp->frame);
        if (pr > prio[P]) prio[P] = pr;
        pop[P]++;
    }
    /* phase 1: probes per population bin */
    uint32_t hist[32] = {0};
    for (uint64_t d = 0; d < nd; d += levels)
        if (canonical[d] == (uint32_t)d && pop[d]) hist[cold_bin(pop[d])]++;
    const uint32_t maxbin = cold_bin(count);
    /* phase 2 in rank-mapped dispatch order: admit nominees */
    uint32_t tickets[32] = {0}, counter = 0;
    for (uint32_t r = 0; r < count; ++r) {
        uint32_t smp = orc_candidate_from_rank(r, rp->width, rp->height, count);
        for (uint32_t v = 0; v < 2; ++v) {
            uint64_t P = ((uint64_t)v * count + smp) * levels;
            if (canonical[P] != (uint32_t)P || prio[P] == 0u) continue;
            uint32_t nsmp = orc_cold_candidate_from_rank(0xffffffffu - prio[P], rp->width, rp->height, count,
                                                         rp->frame);
            uint64_t nom = (uint64_t)nsmp * levels;
            const int32_t* a = keys + 4 * nom;
            const int32_t* b = keys + 4 * P;
            if (canonical[nom] == EMPTY || a[0] != b[0] || a[1] != b[1] || a[2] != b[2] || a[3] != b[3])
                nom = ((uint64_t)count + nsmp) * levels;
            uint32_t bin = cold_bin(pop[P]), higher = 0;
            for (uint32_t hb = bin + 1u; hb <= maxbin; ++hb) higher += hist[hb];
            int admitted = 0;
            if (pop[P] != 0u && higher < budget) admitted = tickets[bin]++ < budget - higher;
            mark[nom] = admitted ? 2u : 0u;
            ordinal[nom] = 0;
            cnt[P] = admitted ? 1u : 0u;
            if (admitted) counter++;
        }
    }
    /* phase 3 in the same order: admitted nominees stay, fair-window
     * receivers take what the budget leaves */
    for (uint32_t r = 0; r < count; ++r) {
        uint32_t smp = orc_candidate_from_rank(r, rp->width, rp->height, count);
        for (uint32_t v = 0; v < 2; ++v) {
            uint64_t d = ((uint64_t)v * count + smp) * levels;
            if (mark[d] != 1u) continue;
            if (orc_candidate_selected(smp, rp) && counter < budget) counter++;
            else mark[d] = 0u;
        }
    }
    for (uint64_t d = 0; d < nd; d += levels)
        if (mark[d] == 1u) ordinal[d] = cnt[canonical[d]]++;
    free(prio);
    free(pop);
}

uint32_t orc_rx_select(uint32_t count, uint32_t levels, const int32_t* keys, const uint32_t* flags,
                       const orc_receiver_params* rp, int cold, uint32_t* canonical, uint32_t* probe_count,
                       uint32_t* probe_offset, uint32_t* sequence, uint32_t* tickets, uint32_t* n_tickets,
                       uint32_t* flags_out) {
    const uint64_t nd = 2ull * count * levels;
    directory D;
    D.cap = orc_rx_directory_capacity(count);
    D.levels = levels;
    D.keys = keys;
    D.slot = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)levels * D.cap);
    for (uint64_t i = 0; i < (uint64_t)levels * D.cap; ++i) D.slot[i] = EMPTY;
    /* per canonical descriptor: count, offset */
    uint32_t* cnt = (uint32_t*)calloc((size_t)nd, sizeof(uint32_t));
    uint32_t* off = (uint32_t*)calloc((size_t)nd, sizeof(uint32_t));
    uint32_t* ordinal = (uint32_t*)calloc((size_t)nd, sizeof(uint32_t));
    /* rank: dedupe in ascending descriptor index, ordinals likewise */
    for (uint64_t d = 0; d < nd; ++d) {
        canonical[d] = EMPTY;
        if (!(flags[d] & ORC_RX_VALID)) continue;
        uint32_t L = (uint32_t)(d % levels);
        canonical[d] = dir_insert(&D, keys + 4 * d, L, (uint32_t)d);
    }
    uint8_t* mark = (uint8_t*)calloc((size_t)nd, 1);
    for (uint64_t d = 0; d < nd; d += levels) mark[d] = canonical[d] != EMPTY && (flags[d] & ORC_RX_RAY);
    if (flags_out) memcpy(flags_out, flags, sizeof(uint32_t) * (size_t)nd);
    if (cold) {
        cold_phases(count, levels, keys, flags, canonical, rp, mark, cnt, ordinal);
        /* level-0 flags after the cold phases: ray = traced, 4 = admitted nominee */
        for (uint64_t d = 0; d < nd; d += levels) {
            if (flags_out)
                flags_out[d] = (flags[d] & ORC_RX_VALID) | (mark[d] ? ORC_RX_RAY : 0u) | (mark[d] == 2u ? 4u : 0u);
            if (mark[d]) mark[d] = 1u;
        }
    } else {
        for (uint64_t d = 0; d < nd; d += levels)
            if (mark[d]) ordinal[d] = cnt[canonical[d]]++;
    }
    /* parents: level L counts from the eight children at L - 1 */
    for (uint32_t L = 1; L < levels; ++L)
        for (uint64_t d = L; d < nd; d += levels) {
            if (canonical[d] != (uint32_t)d) continue;
            const int32_t* k = keys + 4 * d;
            uint32_t lodbits = (uint32_t)k[3] & PROBE_LOD_MASK, sum = 0;
            for (uint32_t c = 0; c < 8; ++c) {
                int32_t ck[4] = {k[0] * 2 + (int32_t)(c & 1u), k[1] * 2 + (int32_t)((c >> 1) & 1u),
                                 k[2] * 2 + (int32_t)((c >> 2) & 1u), (int32_t)(((L - 1u) << 12) | lodbits)};
                uint32_t ci = dir_find(&D, ck, L - 1u);
                if (ci != EMPTY) sum += cnt[ci];
            }
            cnt[d] = sum;
        }
    /* prefix: roots in ascending canonical index, then down the levels */
    uint32_t root = 0;
    for (uint64_t d = levels - 1u; d < nd; d += levels) {
        if (canonical[d] != (uint32_t)d) continue;
        off[d] = root;
        root += cnt[d];
    }
    for (uint32_t L = levels - 1u; L-- > 0;)
        for (uint64_t d = L; d < nd; d += levels) {
            if (canonical[d] != (uint32_t)d) continue;
            const int32_t* k = keys + 4 * d;
            uint32_t lodbits = (uint32_t)k[3] & PROBE_LOD_MASK;
            int32_t parent[4] = {k[0] >> 1, k[1] >> 1, k[2] >> 1, (int32_t)(((L + 1u) << 12) | lodbits)};
            uint32_t pi = dir_find(&D, parent, L + 1u);
            uint32_t o = pi != EMPTY ? off[pi] : 0u;
            uint32_t own = (uint32_t)((k[0] - parent[0] * 2) + (k[1] - parent[1] * 2) * 2 + (k[2] - parent[2] * 2) * 4);
            for (uint32_t sib = 0; sib < own; ++sib) {
                int32_t sk[4] = {parent[0] * 2 + (int32_t)(sib & 1u), parent[1] * 2 + (int32_t)((sib >> 1) & 1u),
                                 parent[2] * 2 + (int32_t)((sib >> 2) & 1u), (int32_t)((L << 12) | lodbits)};
                uint32_t si = dir_find(&D, sk, L);
                if (si != EMPTY) o += cnt[si];
            }
            off[d] = o;
        }
    /* generateDirections: sequences and the ticket list in receiver order */
    uint32_t nt = 0;
    for (uint32_t s = 0; s < count; ++s) {
        int sel = 0;
        for (uint32_t v = 0; v < 2; ++v) {
            uint64_t d = ((uint64_t)v * count + s) * levels;
            sequence[(uint64_t)v * count + s] = EMPTY;
            if (!mark[d]) continue;
            sequence[(uint64_t)v * count + s] = ordinal[d] + off[canonical[d]];
            sel = 1;
        }
        if (sel) tickets[nt++] = s;
    }
    *n_tickets = nt;
    for (uint64_t d = 0; d < nd; ++d) {
        uint32_t c = canonical[d];
        probe_count[d] = c == EMPTY ? 0u : cnt[c];
        probe_offset[d] = c == EMPTY ? EMPTY : off[c];
    }
    free(D.slot);
    free(cnt);
    free(off);
    free(ordinal);
    free(mark);
    return root;
}
