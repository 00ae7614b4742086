/*
 * orc_transport_impl.h -- REAL-generic body of the transport oracle, included
 * twice by orc_transport.c: REAL = float (the product's arithmetic, bit-exact
 * target for the HIP kernels) and REAL = double (the FP64 twin that the f32
 * results are judged against).  K(x) spells a literal in the REAL type.
 */
#ifdef K
#undef K
#endif
#define K(x) ((REAL)(x))

typedef struct {
    REAL v[3];
} SFX(vec);

static inline REAL SFX(dot)(const REAL a[3], const REAL b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
/* min/max/saturate as compare-selects: identical on every IEEE target,
 * including the sign of zero results. */
static inline REAL SFX(smax)(REAL a, REAL b) { return b > a ? b : a; }
static inline REAL SFX(smin)(REAL a, REAL b) { return b < a ? b : a; }
static inline REAL SFX(sat)(REAL x) { return x < K(0.0) ? K(0.0) : (x > K(1.0) ? K(1.0) : x); }
static inline REAL SFX(pow5)(REAL x) {
    REAL x2 = x * x;
    REAL x4 = x2 * x2;
    return x4 * x;
}

/* disneyDiffuseMultiplier (common.slang:299-314); rsqrt as 1/sqrt, pow(x, 5)
 * as repeated products. */
static REAL SFX(disney)(const REAL n[3], const REAL l[3], const REAL v[3], REAL pr) {
    REAL ndl = SFX(sat)(SFX(dot)(n, l));
    REAL ndv = SFX(sat)(SFX(dot)(n, v));
    REAL hs[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
    REAL hl2 = SFX(dot)(hs, hs);
    if (ndl <= K(0.0) || ndv <= K(0.0) || hl2 <= (REAL)1e-20f) return K(0.0);
    REAL rs = K(1.0) / sqrt(hl2);
    REAL h[3] = {hs[0] * rs, hs[1] * rs, hs[2] * rs};
    REAL rough = SFX(smax)(SFX(sat)(pr), (REAL)0.089f);
    rough = rough * rough;
    REAL ldh = SFX(sat)(SFX(dot)(l, h));
    REAL f90 = K(0.5) + ((K(2.0) * rough) * ldh) * ldh;
    REAL ls = K(1.0) + (f90 - K(1.0)) * SFX(pow5)(K(1.0) - ndl);
    REAL vs = K(1.0) + (f90 - K(1.0)) * SFX(pow5)(K(1.0) - ndv);
    return ls * vs;
}

/* outgoingTransportRadianceTracked restricted to the configs' materials
 * (untextured, opaque, black environment, no punctual lights):
 * emission + emissive NEE.  Returns 0 when a visibility query was
 * unsupported (no sample is deposited then). */
static int SFX(shade)(const orc_scene* s, const orc_light* lights, uint32_t nl,
                      const orc_material* mats, uint32_t nm, const orc_gi_params* p,
                      const float o[3], const float d[3], float t, uint32_t tri_slot,
                      uint32_t inst_slot, REAL out[3], orc_dep* dep) {
    const orc_instance* in = &s->inst[inst_slot];
    const float* tri = s->tri + 12ull * tri_slot;
    /* a nonzero instance material (scenes given as scene words) wins */
    uint32_t mi = in->material ? in->material - 1u : fbits(tri[7]);
    if (nm == 0) { out[0] = out[1] = out[2] = K(0.0); return 1; }
    if (mi >= nm) mi = nm - 1;
    const orc_material* mat = &mats[mi];
    REAL rad[3] = {(REAL)mat->emission[0], (REAL)mat->emission[1], (REAL)mat->emission[2]};
    if ((in->flags & 16u) != 0u) goto clamp;
    {
        REAL pos[3], dd[3], gnf[3];
        float gn32[3];
        for (int a = 0; a < 3; ++a) {
            dd[a] = (REAL)d[a];
            pos[a] = (REAL)o[a] + dd[a] * (REAL)t;
        }
        geometric_normal(in, tri, gn32);
        for (int a = 0; a < 3; ++a) gnf[a] = (REAL)gn32[a];
        if (SFX(dot)(gnf, dd) > K(0.0))
            for (int a = 0; a < 3; ++a) gnf[a] = -gnf[a];
        REAL nrm[3] = {gnf[0], gnf[1], gnf[2]};
        REAL md[3] = {-dd[0], -dd[1], -dd[2]};
        if (SFX(dot)(nrm, md) <= K(0.0) || SFX(dot)(nrm, gnf) <= K(0.0)) goto clamp;
        REAL vl = sqrt(SFX(dot)(md, md));
        REAL view[3] = {md[0] / vl, md[1] / vl, md[2] / vl};
        REAL rough = isfinite(mat->roughness) ? (REAL)mat->roughness : K(1.0);
        uint32_t count = p->emissive_samples / 2u;
        if (count < 1u) count = 1u;
        if (nl > 0) {
            REAL bias = (REAL)p->normal_bias;
            REAL sum[3] = {K(0.0), K(0.0), K(0.0)};
            float posf[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
            float sp = p->spacing * 0.05f;
            int32_t cell[3];
            for (int a = 0; a < 3; ++a) cell[a] = (int32_t)floorf(posf[a] / sp);
            uint32_t seed = hash_key(cell[0], cell[1], cell[2], 0, 16777216u);
            float rot[2];
            r2(seed, rot);
            int32_t key[3];
            for (int a = 0; a < 3; ++a) key[a] = (int32_t)floorf(posf[a] / p->spacing);
            uint32_t sel_seed = hash_key(key[0], key[1], key[2], 0, 65536u);
            uint32_t frame = p->frame ^ 0xE11E0001u;
            float ra[2], rb[2];
            r2(frame, ra);
            r2(sel_seed, rb);
            float phase = fracf_(ra[0] + rb[0]);
            for (uint32_t smp = 0; smp < count; ++smp) {
                float q[2], uv[2];
                r2(smp + frame * count + 1u, q);
                uv[0] = fracf_(q[0] + rot[0]);
                uv[1] = fracf_(q[1] + rot[1]);
                float selection = fracf_(((float)smp + 0.5f) / (float)count + phase);
                uint32_t lo = 0, hi = nl - 1u;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) / 2u;
                    if (lights[mid].direction[3] < selection) lo = mid + 1u;
                    else hi = mid;
                }
                const orc_light* L = &lights[lo];
                REAL lp[3] = {(REAL)L->position[0], (REAL)L->position[1], (REAL)L->position[2]};
                REAL ldir[3] = {(REAL)L->direction[0], (REAL)L->direction[1], (REAL)L->direction[2]};
                REAL rel[3] = {pos[0] - lp[0], pos[1] - lp[1], pos[2] - lp[2]};
                if (L->edge1[3] <= 0.0f ||
                    (L->position[3] != -2.0f && SFX(dot)(ldir, rel) <= K(0.0)))
                    continue;
                REAL root = sqrt((REAL)uv[0]);
                REAL wa = root * (K(1.0) - (REAL)uv[1]);
                REAL wb = root * (REAL)uv[1];
                REAL target[3], tl[3];
                for (int a = 0; a < 3; ++a) {
                    target[a] = (lp[a] + (REAL)L->edge1[a] * wa) + (REAL)L->edge2[a] * wb;
                    tl[a] = target[a] - pos[a];
                }
                REAL ldist = sqrt(SFX(dot)(tl, tl));
                if (ldist <= bias * K(0.25)) continue;
                REAL Ld[3] = {tl[0] / ldist, tl[1] / ldist, tl[2] / ldist};
                REAL cosine = SFX(smax)(K(0.0), SFX(dot)(nrm, Ld));
                REAL mL[3] = {-Ld[0], -Ld[1], -Ld[2]};
                REAL ec = SFX(dot)(ldir, mL);
                if (L->position[3] == -2.0f) ec = fabs(ec);
                REAL ecos = SFX(smax)(K(0.0), ec);
                if (cosine <= K(0.0) || ecos <= K(0.0) || SFX(dot)(gnf, Ld) <= K(0.0)) continue;
                REAL em[3] = {(REAL)L->color[0], (REAL)L->color[1], (REAL)L->color[2]};
                if (!(em[0] > K(0.0) || em[1] > K(0.0) || em[2] > K(0.0))) continue;
                REAL pdf = (((REAL)L->edge1[3] * ldist) * ldist) / ((REAL)L->color[3] * ecos);
                REAL plane = fabs(SFX(dot)(ldir, rel));
                REAL ob = SFX(smin)(bias, plane * K(0.25));
                float pos32[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
                float gn3[3] = {(float)gnf[0], (float)gnf[1], (float)gnf[2]};
                float origin[3];
                ray_offset(pos32, gn3, (float)ob, origin);
                REAL delta[3] = {target[0] - (REAL)origin[0], target[1] - (REAL)origin[1],
                                 target[2] - (REAL)origin[2]};
                REAL dist = sqrt(SFX(dot)(delta, delta));
                if (dist <= bias * K(0.25)) continue;
                REAL dir[3] = {delta[0] / dist, delta[1] / dist, delta[2] / dist};
                REAL lim = dist - SFX(smin)(bias * K(0.25), dist * (REAL)0.001f);
                orc_ray ray;
                ray.ox = origin[0]; ray.oy = origin[1]; ray.oz = origin[2];
                ray.tmax = (float)lim;
                ray.dx = (float)dir[0]; ray.dy = (float)dir[1]; ray.dz = (float)dir[2];
                ray.flags = ORC_RAY_ANY_HIT;
                orc_hit vh;
                orc_trace_ray(s, &ray, &vh, NULL, NULL, NULL);
                if (dep) {
                    const float ro[3] = {ray.ox, ray.oy, ray.oz}, rd[3] = {ray.dx, ray.dy, ray.dz};
                    dep_ray(dep, s, ro, rd, ray.tmax, 0, 0.0f);
                }
                if (vh.flags & ORC_HIT_OCCLUDED) continue;
                if (vh.flags & ORC_ERR_UNSUPPORTED) return 0;
                REAL dm = SFX(disney)(nrm, Ld, view, rough) / (REAL)ORC_PI_F;
                for (int c = 0; c < 3; ++c) {
                    REAL response = (REAL)mat->albedo[c] * dm;
                    sum[c] = sum[c] + em[c] * ((response * cosine) / SFX(smax)(pdf, (REAL)1e-20f));
                }
            }
            for (int c = 0; c < 3; ++c) rad[c] = rad[c] + sum[c] / (REAL)count;
        }
    }
clamp:
    if (!(isfinite(rad[0]) && isfinite(rad[1]) && isfinite(rad[2]))) {
        out[0] = out[1] = out[2] = K(0.0);
        return 1;
    }
    for (int c = 0; c < 3; ++c) out[c] = SFX(smax)(rad[c], K(0.0));
    return 1;
}

typedef struct {
    const orc_scene* s;
    const orc_light* lights;
    uint32_t nl;
    const orc_material* mats;
    uint32_t nm;
    const orc_gi_params* p;
    REAL* intervals;
    uint32_t* samples; /* nullable: supported rays per interval */
    uint64_t begin, end; /* positions k of the traced intervals g = k * stride */
    uint64_t stride;
    float* dep; /* nullable (f32 only): 8 floats per interval, this trace's dependency */
} SFX(trace_job);

static void* SFX(trace_worker)(void* arg) {
    SFX(trace_job)* j = (SFX(trace_job)*)arg;
    const orc_gi_params* p = j->p;
    uint32_t levels = p->levels;
    uint64_t off[17];
    for (uint32_t l = 0; l <= levels; ++l) off[l] = orc_gi_level_offset(p, l);
    for (uint64_t kpos = j->begin; kpos < j->end; ++kpos) {
        const uint64_t g = kpos * j->stride;
        uint32_t L = 0;
        while (L + 1 < levels && g >= off[L + 1]) ++L;
        uint64_t local = g - off[L];
        uint32_t dims[3];
        orc_gi_level_dims(p, L, dims);
        uint32_t D = orc_gi_level_directions(p, L);
        uint32_t nphi = p->base_phi << L, ny = p->base_y << L;
        uint32_t k = (uint32_t)(local % D);
        uint64_t probe = local / D;
        uint32_t ix = (uint32_t)(probe % dims[0]);
        uint32_t iy = (uint32_t)((probe / dims[0]) % dims[1]);
        uint32_t iz = (uint32_t)(probe / ((uint64_t)dims[0] * dims[1]));
        float sl = p->spacing * ldexpf(1.0f, (int)L);
        float c[3] = {p->origin[0] + ((float)ix + 0.5f) * sl, p->origin[1] + ((float)iy + 0.5f) * sl,
                      p->origin[2] + ((float)iz + 0.5f) * sl};
        float start, end;
        orc_gi_interval_bounds(p, L, &start, &end);
        int terminal = p->terminal_tail && L + 1 == levels;
        float limit = terminal ? 1e20f : end - start;
        REAL jsum[3] = {K(0.0), K(0.0), K(0.0)}, bsum = K(0.0);
        uint32_t n = 0;
        orc_dep idep = dep_empty();
        uint32_t R = p->rays_per_interval ? p->rays_per_interval : 1u;
        for (uint32_t r = 0; r < R; ++r) {
            float ouv[2], d[3];
            ray_offset_uv(p->seed, r, ouv);
            orc_bin_direction(k, nphi, ny, ouv[0], ouv[1], d);
            orc_ray ray;
            ray.ox = c[0] + d[0] * start;
            ray.oy = c[1] + d[1] * start;
            ray.oz = c[2] + d[2] * start;
            ray.tmax = limit;
            ray.dx = d[0]; ray.dy = d[1]; ray.dz = d[2];
            ray.flags = 0;
            orc_hit h;
            uint32_t ts, is;
            orc_trace_ray(j->s, &ray, &h, NULL, &ts, &is);
            if (h.flags & ORC_ERR_UNSUPPORTED) continue;
            REAL rad[3] = {K(0.0), K(0.0), K(0.0)};
            REAL beta;
            /* the ray's segment, then its NEE occlusion segments */
            orc_dep rdep = dep_empty();
            float o3[3] = {ray.ox, ray.oy, ray.oz};
            if (j->dep) dep_ray(&rdep, j->s, o3, d, ray.tmax, h.triangle != ORC_EMPTY, h.t);
            if (h.triangle != ORC_EMPTY) {
                if (!SFX(shade)(j->s, j->lights, j->nl, j->mats, j->nm, p, o3, d, h.t, ts, is, rad,
                                j->dep ? &rdep : NULL))
                    continue;
                beta = K(0.0);
            } else {
                beta = terminal ? K(0.0) : K(1.0); /* escape: black environment */
            }
            if (j->dep) dep_union(&idep, &rdep);
            for (int a = 0; a < 3; ++a) jsum[a] = jsum[a] + rad[a];
            bsum = bsum + beta;
            n++;
        }
        if (j->samples) j->samples[g] = n;
        if (j->dep) dep_store(j->dep + 8 * g, &idep);
        REAL* out = j->intervals + 4 * g;
        if (n == 0) {
            out[0] = out[1] = out[2] = K(0.0);
            out[3] = K(-1.0);
        } else {
            REAL fn = (REAL)n;
            out[0] = jsum[0] / fn; out[1] = jsum[1] / fn; out[2] = jsum[2] / fn;
            out[3] = bsum / fn;
        }
    }
    return NULL;
}

static void SFX(gi_trace)(const orc_scene* s, const orc_light* lights, uint32_t nl,
                          const orc_material* mats, uint32_t nm, const orc_gi_params* p,
                          REAL* intervals, uint32_t* samples, int threads, uint64_t stride, float* dep) {
    if (stride < 1) stride = 1;
    uint64_t total = (orc_gi_level_offset(p, p->levels) + stride - 1) / stride;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    SFX(trace_job) jobs[64];
    pthread_t tids[64];
    uint64_t per = (total + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].s = s; jobs[t].lights = lights; jobs[t].nl = nl; jobs[t].mats = mats;
        jobs[t].nm = nm; jobs[t].p = p; jobs[t].intervals = intervals; jobs[t].samples = samples;
        jobs[t].begin = per * (uint64_t)t < total ? per * (uint64_t)t : total;
        jobs[t].end = jobs[t].begin + per < total ? jobs[t].begin + per : total;
        jobs[t].stride = stride;
        jobs[t].dep = dep;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&tids[t], NULL, SFX(trace_worker), &jobs[t]);
    SFX(trace_worker)(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(tids[t], NULL);
}

/* Support-weighted trilinear child lookup (interpolateMergedChildSources,
 * split.slang:389-479) at level Lc for direction `dir` of probe-relative
 * position rel (world position minus grid origin). */
static void SFX(interp_child)(const orc_gi_params* p, const REAL* merged_c, uint32_t Lc,
                              uint32_t dir, const float rel[3], REAL value[4]) {
    uint32_t dims[3];
    orc_gi_level_dims(p, Lc, dims);
    uint32_t D = orc_gi_level_directions(p, Lc);
    float sc = p->spacing * ldexpf(1.0f, (int)Lc);
    float grid[3], fr[3];
    int32_t base[3];
    for (int a = 0; a < 3; ++a) {
        grid[a] = rel[a] / sc - 0.5f;
        float fl = floorf(grid[a]);
        base[a] = (int32_t)fl;
        fr[a] = grid[a] - fl;
    }
    REAL sum[3] = {K(0.0), K(0.0), K(0.0)}, total = K(0.0), potential = K(0.0);
    for (uint32_t corner = 0; corner < 8u; ++corner) {
        uint32_t bx = corner & 1u, by = (corner >> 1) & 1u, bz = corner >> 2;
        REAL ax = bx == 0u ? K(1.0) - (REAL)fr[0] : (REAL)fr[0];
        REAL ay = by == 0u ? K(1.0) - (REAL)fr[1] : (REAL)fr[1];
        REAL az = bz == 0u ? K(1.0) - (REAL)fr[2] : (REAL)fr[2];
        REAL w = (ax * ay) * az;
        if (!(w > K(0.0))) continue;
        potential = potential + w;
        int64_t cx = (int64_t)base[0] + bx, cy = (int64_t)base[1] + by, cz = (int64_t)base[2] + bz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= dims[0] || cy >= dims[1] || cz >= dims[2]) continue;
        uint64_t probe = ((uint64_t)cz * dims[1] + (uint64_t)cy) * dims[0] + (uint64_t)cx;
        const REAL* smp = merged_c + 4 * (probe * D + dir);
        if (!(smp[3] > K(0.0)) || !(isfinite(smp[0]) && isfinite(smp[1]) && isfinite(smp[2]) &&
                                    isfinite(smp[3])))
            continue;
        REAL sw = w * SFX(sat)(smp[3]);
        for (int c = 0; c < 3; ++c) sum[c] = sum[c] + smp[c] * sw;
        total = total + sw;
    }
    if (total > K(0.0)) {
        for (int c = 0; c < 3; ++c) value[c] = sum[c] / total;
        value[3] = SFX(sat)(total / SFX(smax)(potential, (REAL)1e-8f));
    } else {
        value[0] = value[1] = value[2] = value[3] = K(0.0);
    }
}

/* mergeEntry for one history-free update, top level down. */
static void SFX(gi_merge)(const orc_gi_params* p, const REAL* intervals, const float* history, REAL* merged) {
    for (int Li = (int)p->levels - 1; Li >= 0; --Li) {
        uint32_t L = (uint32_t)Li;
        uint32_t dims[3];
        uint32_t P = orc_gi_level_dims(p, L, dims);
        uint32_t D = orc_gi_level_directions(p, L);
        uint64_t off = orc_gi_level_offset(p, L);
        uint64_t offc = orc_gi_level_offset(p, L + 1);
        uint32_t nphi = p->base_phi << L;
        float sl = p->spacing * ldexpf(1.0f, (int)L);
        for (uint32_t probe = 0; probe < P; ++probe) {
            uint32_t ix = probe % dims[0], iy = (probe / dims[0]) % dims[1], iz = probe / (dims[0] * dims[1]);
            float rel[3] = {((float)ix + 0.5f) * sl, ((float)iy + 0.5f) * sl, ((float)iz + 0.5f) * sl};
            for (uint32_t k = 0; k < D; ++k) {
                uint64_t g = off + (uint64_t)probe * D + k;
                const REAL* iv = intervals + 4 * g;
                REAL* out = merged + 4 * g;
                /* own radiance needs both masses, children only transmittance
                 * (beginMergeGroup, split.slang:1533-1545); with no history
                 * both masses are 1 for a sampled interval, 0 otherwise */
                int has;
                REAL tr;
                if (history) {
                    float rm = history[16 * g + 3], tm = history[16 * g + 5];
                    tr = tm > 0.0f ? SFX(sat)(iv[3]) : K(0.0);
                    has = rm > 0.0f && tm > 0.0f;
                } else {
                    has = iv[3] >= K(0.0);
                    tr = has ? SFX(sat)(iv[3]) : K(0.0);
                }
                if (!has && !(tr > K(0.0))) {
                    out[0] = out[1] = out[2] = out[3] = K(0.0);
                    continue;
                }
                REAL support = has ? K(1.0) - tr : K(0.0);
                REAL rad[3] = {K(0.0), K(0.0), K(0.0)};
                if (support > K(0.0)) { rad[0] = iv[0]; rad[1] = iv[1]; rad[2] = iv[2]; }
                REAL csum[3] = {K(0.0), K(0.0), K(0.0)}, csup = K(0.0);
                if (tr > K(0.0) && L + 1 < p->levels) {
                    uint32_t x = k % nphi, y = k / nphi;
                    for (uint32_t child = 0; child < 4u; ++child) {
                        uint32_t fx = x * 2u + (child & 1u), fy = y * 2u + (child >> 1);
                        uint32_t fd = fx + fy * (2u * nphi);
                        REAL v[4];
                        SFX(interp_child)(p, merged + 4 * offc, L + 1, fd, rel, v);
                        if (!(v[3] > K(0.0))) continue;
                        for (int c = 0; c < 3; ++c) csum[c] = csum[c] + v[c] * v[3];
                        csup = csup + v[3];
                    }
                }
                if (csup > K(0.0)) {
                    for (int c = 0; c < 3; ++c) rad[c] = rad[c] + (tr * csum[c]) * K(0.25);
                    support = support + (tr * csup) * K(0.25);
                }
                if (support > K(0.0)) {
                    for (int c = 0; c < 3; ++c) out[c] = rad[c] / support;
                    out[3] = SFX(sat)(support);
                } else {
                    out[0] = out[1] = out[2] = out[3] = K(0.0);
                }
            }
        }
    }
}

/* decodeNormal (common.slang:370-375). */
static void SFX(decode_normal)(float px, float py, REAL out[3]) {
    float x = px * 2.0f - 1.0f, y = py * 2.0f - 1.0f;
    float z = 1.0f - fabsf(x) - fabsf(y);
    float nz = -z;
    float t = nz < 0.0f ? 0.0f : (nz > 1.0f ? 1.0f : nz);
    x = x + (x >= 0.0f ? -t : t);
    y = y + (y >= 0.0f ? -t : t);
    float l = sqrtf((x * x + y * y) + z * z);
    out[0] = (REAL)(x / l); out[1] = (REAL)(y / l); out[2] = (REAL)(z / l);
}

/* packProbes (split.slang:1788-1918) for every level-0 probe. */
static void SFX(gi_pack)(const orc_gi_params* p, const REAL* merged, REAL* probes) {
    uint32_t dims[3];
    uint32_t P = orc_gi_level_dims(p, 0, dims);
    uint32_t D = orc_gi_level_directions(p, 0);
    REAL (*dirs)[3] = malloc(sizeof(REAL) * 3 * D);
    for (uint32_t k = 0; k < D; ++k) {
        float d[3];
        orc_bin_direction(k, p->base_phi, p->base_y, 0.5f, 0.5f, d);
        dirs[k][0] = (REAL)d[0]; dirs[k][1] = (REAL)d[1]; dirs[k][2] = (REAL)d[2];
    }
    REAL norms[36][3];
    for (uint32_t t = 0; t < 36; ++t)
        SFX(decode_normal)(((float)(t % 6u) + 0.5f) / 6.0f, ((float)(t / 6u) + 0.5f) / 6.0f, norms[t]);
    REAL scale_k = K(4.0) / (REAL)D;
    for (uint32_t probe = 0; probe < P; ++probe) {
        const REAL* m = merged + 4ull * probe * D;
        for (uint32_t t = 0; t < 36; ++t) {
            REAL sum[3] = {K(0.0), K(0.0), K(0.0)}, nominal = K(0.0), observed = K(0.0);
            for (uint32_t k = 0; k < D; ++k) {
                REAL c = SFX(smax)(SFX(dot)(norms[t], dirs[k]), K(0.0));
                nominal = nominal + c;
                if (c <= K(0.0)) continue;
                const REAL* smp = m + 4 * k;
                int ok = smp[3] > K(0.0) && isfinite(smp[0]) && isfinite(smp[1]) &&
                         isfinite(smp[2]) && isfinite(smp[3]);
                if (!ok) continue;
                REAL w = c * SFX(sat)(smp[3]);
                for (int q = 0; q < 3; ++q) sum[q] = sum[q] + smp[q] * w;
                observed = observed + w;
            }
            REAL* out = probes + 4 * (36ull * probe + t);
            out[0] = out[1] = out[2] = out[3] = K(0.0);
            if (nominal > K(0.0) && observed > K(0.0)) {
                REAL sc = nominal * scale_k;
                for (int q = 0; q < 3; ++q) out[q] = (sum[q] / observed) * sc;
                out[3] = SFX(sat)(observed / nominal);
            }
        }
    }
    free(dirs);
}

static void SFX(wrap_texel)(int32_t* x, int32_t* y) {
    int32_t px = *x, py = *y;
    if (px < 0) { int32_t nx = -px - 1, ny = 5 - py; px = nx; py = ny; }
    else if (px >= 6) { int32_t nx = 11 - px, ny = 5 - py; px = nx; py = ny; }
    if (py < 0) { int32_t nx = 5 - px, ny = -py - 1; px = nx; py = ny; }
    else if (py >= 6) { int32_t nx = 5 - px, ny = 11 - py; px = nx; py = ny; }
    *x = px; *y = py;
}

/* irradianceLookup (world_resolve.slang:25-40): the octahedral footprint of
 * the unit normal -- 4 texels and bilinear weights. */
static void SFX(resolve_lookup)(float nx, float ny, float nz, uint32_t tex[4], REAL bw[4]) {
    float sabs = (fabsf(nx) + fabsf(ny)) + fabsf(nz);
    float ex = nx / sabs, ey = ny / sabs, ez = nz / sabs;
    if (ez < 0.0f) {
        float tx = (1.0f - fabsf(ey)) * (ex >= 0.0f ? 1.0f : -1.0f);
        float ty = (1.0f - fabsf(ex)) * (ey >= 0.0f ? 1.0f : -1.0f);
        ex = tx; ey = ty;
    }
    float gx = (ex * 0.5f + 0.5f) * 6.0f - 0.5f, gy = (ey * 0.5f + 0.5f) * 6.0f - 0.5f;
    float flx = floorf(gx), fly = floorf(gy);
    float fx = gx - flx, fy = gy - fly;
    float cx = 1.0f - fx, cy = 1.0f - fy;
    bw[0] = (REAL)(cx * cy); bw[1] = (REAL)(fx * cy); bw[2] = (REAL)(cx * fy); bw[3] = (REAL)(fx * fy);
    int32_t bxi = (int32_t)flx, byi = (int32_t)fly;
    for (int c = 0; c < 4; ++c) {
        int32_t tx = bxi + (c & 1), ty = byi + (c >> 1);
        SFX(wrap_texel)(&tx, &ty);
        tex[c] = (uint32_t)(ty * 6 + tx);
    }
}

/* diffuseAtLod + diffuseProbePage (world_resolve.slang:52-170) over one dense
 * probe grid: 8 probes trilinear x 4 texels bilinear, support-normalised.
 * Probes outside the grid are missing (they count in the potential weight).
 * probes == NULL: every probe missing. */
static void SFX(diffuse_at)(const float origin[3], float spacing, const uint32_t dims[3], const REAL* probes,
                            const float* pos, const uint32_t tex[4], const REAL bw[4], REAL val[3],
                            REAL* cov) {
    float rel[3] = {pos[0] - origin[0], pos[1] - origin[1], pos[2] - origin[2]};
    float fr[3];
    int32_t base[3];
    for (int a = 0; a < 3; ++a) {
        float g = rel[a] / spacing - 0.5f;
        float fl = floorf(g);
        base[a] = (int32_t)fl;
        fr[a] = g - fl;
    }
    REAL sum[3] = {K(0.0), K(0.0), K(0.0)}, observed = K(0.0), potential = K(0.0);
    for (uint32_t corner = 0; corner < 8u; ++corner) {
        uint32_t bx = corner & 1u, by = (corner >> 1) & 1u, bz = corner >> 2;
        REAL ax = bx == 0u ? K(1.0) - (REAL)fr[0] : (REAL)fr[0];
        REAL ay = by == 0u ? K(1.0) - (REAL)fr[1] : (REAL)fr[1];
        REAL az = bz == 0u ? K(1.0) - (REAL)fr[2] : (REAL)fr[2];
        REAL w = (ax * ay) * az;
        if (!(w > K(0.0))) continue;
        potential = potential + w;
        if (!probes) continue;
        int64_t qx = (int64_t)base[0] + bx, qy = (int64_t)base[1] + by, qz = (int64_t)base[2] + bz;
        if (qx < 0 || qy < 0 || qz < 0 || qx >= dims[0] || qy >= dims[1] || qz >= dims[2]) continue;
        uint64_t probe = ((uint64_t)qz * dims[1] + (uint64_t)qy) * dims[0] + (uint64_t)qx;
        /* diffuseProbePage */
        REAL psum[3] = {K(0.0), K(0.0), K(0.0)}, pobs = K(0.0);
        for (int c = 0; c < 4; ++c) {
            if (!(bw[c] > K(0.0))) continue;
            const REAL* smp = probes + 4 * (36ull * probe + tex[c]);
            if (!(smp[3] > K(0.0)) || !(isfinite(smp[0]) && isfinite(smp[1]) &&
                                        isfinite(smp[2]) && isfinite(smp[3])))
                continue;
            REAL sw = bw[c] * SFX(sat)(smp[3]);
            for (int q = 0; q < 3; ++q) psum[q] = psum[q] + smp[q] * sw;
            pobs = pobs + sw;
        }
        REAL pv[3] = {K(0.0), K(0.0), K(0.0)};
        if (pobs > K(0.0))
            for (int q = 0; q < 3; ++q) pv[q] = psum[q] / pobs;
        REAL support = SFX(sat)(pobs);
        if (!(support > K(0.0))) continue;
        REAL sw = w * support;
        for (int q = 0; q < 3; ++q) sum[q] = sum[q] + pv[q] * sw;
        observed = observed + sw;
    }
    val[0] = val[1] = val[2] = K(0.0);
    *cov = K(0.0);
    if (potential > K(0.0)) {
        *cov = SFX(sat)(observed / potential);
        if (observed > K(0.0))
            for (int q = 0; q < 3; ++q) val[q] = sum[q] / observed;
    }
}

/* resolveLodSource (world_resolve.slang:172-182) and the raw alpha code
 * (src/lib.rs:320-335): 0 not eligible, -1 unsupported, 1 + coverage / 4. */
static void SFX(resolve_combine)(const REAL lval[3], REAL lcov, const REAL hval[3], REAL hcov, float blend,
                                 REAL* o) {
    REAL lw = lcov * (K(1.0) - (REAL)blend), hw = hcov * (REAL)blend, tw = lw + hw;
    REAL coverage = SFX(sat)(tw);
    REAL res[3] = {K(0.0), K(0.0), K(0.0)};
    if (tw > K(0.0))
        for (int q = 0; q < 3; ++q) res[q] = (lval[q] * lw + hval[q] * hw) / tw;
    int supported = coverage > K(0.0) && isfinite(res[0]) && isfinite(res[1]) && isfinite(res[2]);
    o[0] = res[0]; o[1] = res[1]; o[2] = res[2];
    o[3] = supported ? K(1.0) + K(0.25) * coverage : K(-1.0);
}

/* Receiver eligibility (finite normal, |n|^2 > 0.01) and its unit normal. */
static int SFX(resolve_normal)(const float* r, float* nx, float* ny, float* nz) {
    float x = r[3], y = r[4], z = r[5];
    float nn = (x * x + y * y) + z * z;
    if (!(finitef_(x) && finitef_(y) && finitef_(z)) || !(nn > 0.01f)) return 0;
    float nl = sqrtf(nn);
    *nx = x / nl; *ny = y / nl; *nz = z / nl;
    return 1;
}

/* worldResolve at LOD 0 (world_resolve.slang:1-196,281-313): the LOD pair
 * with blend 0 (no high LOD, its coverage 1). */
static void SFX(world_resolve)(const orc_gi_params* p, const REAL* probes, const float* recv,
                               uint32_t n, REAL* out) {
    uint32_t dims[3];
    orc_gi_level_dims(p, 0, dims);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = recv + 6ull * i;
        REAL* o = out + 4ull * i;
        o[0] = o[1] = o[2] = o[3] = K(0.0);
        float nx, ny, nz;
        if (!SFX(resolve_normal)(r, &nx, &ny, &nz)) continue; /* not eligible */
        uint32_t tex[4];
        REAL bw[4];
        SFX(resolve_lookup)(nx, ny, nz, tex, bw);
        REAL val[3], cov;
        SFX(diffuse_at)(p->origin, p->spacing, dims, probes, r, tex, bw, val, &cov);
        const REAL zero[3] = {K(0.0), K(0.0), K(0.0)};
        SFX(resolve_combine)(val, cov, zero, K(1.0), 0.0f, o);
    }
}

/* resolveSplitDiffuse (world_resolve.slang:184-219) over a pyramid of dense
 * probe grids, grid k at LOD k: the LOD pair and blend of each receiver
 * (orc_lod_select), each LOD's diffuse lookup (LODs >= count have no probes),
 * resolveLodSource. */
static void SFX(world_resolve_lods)(const orc_gi_params* const* grids, const REAL* const* probes, uint32_t count,
                                    const orc_lod_view* view, const float* recv, uint32_t n, REAL* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = recv + 6ull * i;
        REAL* o = out + 4ull * i;
        o[0] = o[1] = o[2] = o[3] = K(0.0);
        float nx, ny, nz;
        if (!SFX(resolve_normal)(r, &nx, &ny, &nz)) continue;
        uint32_t tex[4];
        REAL bw[4];
        SFX(resolve_lookup)(nx, ny, nz, tex, bw);
        uint32_t low, high;
        float blend;
        orc_lod_select(view, r, &low, &high, &blend);
        REAL lval[3] = {K(0.0), K(0.0), K(0.0)}, lcov = K(0.0);
        REAL hval[3] = {K(0.0), K(0.0), K(0.0)}, hcov = K(1.0);
        uint32_t dims[3];
        if (low < count) {
            orc_gi_level_dims(grids[low], 0, dims);
            SFX(diffuse_at)(grids[low]->origin, grids[low]->spacing, dims, probes[low], r, tex, bw, lval, &lcov);
        }
        if (high != low) {
            hcov = K(0.0);
            if (high < count) {
                orc_gi_level_dims(grids[high], 0, dims);
                SFX(diffuse_at)(grids[high]->origin, grids[high]->spacing, dims, probes[high], r, tex, bw, hval,
                                &hcov);
            }
        }
        SFX(resolve_combine)(lval, lcov, hval, hcov, blend, o);
    }
}
#undef K
