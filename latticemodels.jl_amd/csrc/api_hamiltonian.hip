// C ABI of liblmhip.so, the Hamiltonian: geometry, on-site energies, fields,
// the CSC boundary, value classes, and the solver update (the Taylor-order
// rule, its result kept on the device or read back).

#include "api_internal.h"

using namespace lmhapi;

namespace lmhapi {
void free_dict(lmh::DevHam &h) {
    dfree(h.dpack);
    dfree(h.drec);
    dfree(h.dcls);
    dfree(h.drep);
    dfree(h.dval);
    dfree(h.doff);
    h.ndict = h.noff = h.dict_ok = 0;
}

// H's value classes for the fused passes (zmax <= 4, n <= 10240): every slot
// is evaluated for `dict_probes` sets of field amplitudes (the field's kinds
// and points kept; probe 0 has all amplitudes zero), and slots whose values
// agree bit for bit at every probe -- directly or conjugated -- share a class.
// Each H build re-checks the classes against the actual values (red[2]); a
// mismatch makes the fused passes read val instead, so the classes change
// speed, never results.  Column offsets (column - site) must take at most 16
// values and the classes at most 2048.
std::vector<double> dict_signature(const lmh_context *ctx) {
    std::vector<double> sig;
    for (const auto &f : ctx->field_host) {
        sig.push_back((double)f.kind);
        sig.push_back(f.px);
        sig.push_back(f.py);
    }
    return sig;
}

hipError_t build_dict(lmh_context *ctx) {
    lmh::DevHam &h = ctx->h;
    free_dict(h);
    ctx->dict_geom = ctx->geom_version;
    ctx->dict_probes_built = ctx->dict_probes;
    ctx->dict_sig = dict_signature(ctx);
    ctx->built_version = 0;                  // the next update rebuilds H with the classes
    const int64_t n = h.n, z = h.zmax, nslots = z * n;
    if (ctx->dict_probes <= 0 || !lmh::fused_supported(h) || z == 0 || z > 4) return hipSuccess;
    const int np = ctx->dict_probes;
    const int64_t nf = h.nfield;
    std::vector<lmh::FieldComp> pf((size_t)np * nf);
    for (int p = 0; p < np; ++p)
        for (int64_t c = 0; c < nf; ++c) {
            lmh::FieldComp f = ctx->field_host[c];
            f.a = p == 0 ? 0.0 : (p == 1 ? 0.3183098861837907 * (1.0 + 0.37 * c) : 0.1414213562373095 * (1.0 + 0.61 * c));
            pf[(size_t)p * nf + c] = f;
        }
    lmh::FieldComp *dpf = nullptr;
    double2 *dprobe = nullptr;
    std::vector<double2> probe((size_t)np * nslots);
    std::vector<int32_t> col(nslots);
    hipError_t e = hipSuccess;
    if (nf) e = dalloc(&dpf, pf.size());
    if (e == hipSuccess) e = dalloc(&dprobe, probe.size());
    if (e == hipSuccess && nf) e = hipMemcpyAsync(dpf, pf.data(), pf.size() * sizeof(lmh::FieldComp), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = lmh::launch_slot_probe(h, dpf, np, dprobe, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(probe.data(), dprobe, probe.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(col.data(), h.col, 4 * nslots, hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    dfree(dpf);
    dfree(dprobe);
    if (e != hipSuccess) return e;
    struct Key {
        uint64_t w[6];
        bool operator==(const Key &o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            uint64_t x = 0x9e3779b97f4a7c15ull;
            for (uint64_t v : k.w) x = (x ^ v) * 0x100000001b3ull;
            return (size_t)(x ^ (x >> 29));
        }
    };
    auto bits = [](double x) {
        x += 0.0;                              // -0.0 and 0.0 are one value
        uint64_t u;
        std::memcpy(&u, &x, 8);
        return u;
    };
    std::unordered_map<Key, int32_t, KeyHash> classes;
    std::vector<int32_t> rep{-1};            // class 0: the value 0 of empty slots
    std::vector<int32_t> cls(nslots, 0);
    for (int64_t q = 0; q < nslots; ++q) {
        if (col[q] < 0) continue;
        Key k{}, kc{};
        for (int p = 0; p < np && p < 3; ++p) {
            const double2 v = probe[(size_t)p * nslots + q];
            k.w[2 * p] = kc.w[2 * p] = bits(v.x);
            k.w[2 * p + 1] = bits(v.y);
            kc.w[2 * p + 1] = bits(-v.y);
        }
        auto it = classes.find(k);
        if (it != classes.end()) {
            cls[q] = it->second;
        } else if ((it = classes.find(kc)) != classes.end()) {
            cls[q] = it->second | (1 << 15);
        } else {
            if ((int)rep.size() >= lmh::kDictMaxClasses) return hipSuccess;   // too many: no classes
            classes.emplace(k, (int32_t)rep.size());
            cls[q] = (int32_t)rep.size();
            rep.push_back((int32_t)q);
        }
    }
    std::vector<int32_t> offs{0};
    std::vector<uint64_t> pack(n, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t u = 0; u < z; ++u) {
            const int64_t q = u * n + i;
            if (col[q] < 0) continue;
            const int32_t delta = col[q] - (int32_t)i;
            int oid = (int)(std::find(offs.begin(), offs.end(), delta) - offs.begin());
            if (oid == (int)offs.size()) {
                if ((int)offs.size() >= lmh::kDictMaxOffsets) return hipSuccess;   // no classes
                offs.push_back(delta);
            }
            const uint64_t sl = (uint64_t)oid | (uint64_t)(cls[q] & 0x7ff) << 4 | (uint64_t)((cls[q] >> 15) & 1) << 15;
            pack[i] |= sl << (16 * u);
        }
    const int noff = (int)offs.size();
    offs.resize(lmh::kDictMaxOffsets, 0);
    if (e == hipSuccess) e = dalloc(&h.dpack, n);
    if (e == hipSuccess) e = dalloc(&h.drec, n);
    if (e == hipSuccess) e = dalloc(&h.dcls, nslots);
    if (e == hipSuccess) e = dalloc(&h.drep, rep.size());
    if (e == hipSuccess) e = dalloc(&h.dval, rep.size());
    if (e == hipSuccess) e = dalloc(&h.doff, offs.size());
    if (e == hipSuccess) e = hipMemcpy(h.dpack, pack.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.dcls, cls.data(), 4 * nslots, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.drep, rep.data(), 4 * rep.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.doff, offs.data(), 4 * offs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_dict(h);
        return e;
    }
    h.ndict = (int)rep.size();
    h.noff = noff;
    return hipSuccess;
}

void free_geometry(lmh::DevHam &h) {
    dfree(h.ri);
    dfree(h.rj);
    dfree(h.amp);
    dfree(h.twist);
    dfree(h.bslot);
    dfree(h.bi);
    dfree(h.bj);
    dfree(h.cptr);
    dfree(h.cbond);
    dfree(h.col);
    dfree(h.col4);
    dfree(h.val);
    dfree(h.diag);
    dfree(h.rec);
    dfree(h.onsite);
    dfree(h.gcols);
    dfree(h.ucols);
    dfree(h.uptr);
    dfree(h.umap);
    h.ngroups = h.umax = 0;
    h.group = 0;
    dfree(h.sslot);
    dfree(h.sdir);
    dfree(h.stval);
    dfree(h.sdiag);
    h.sg = lmh::StripGeom{};
    h.strip_ok = h.strip_xreal = 0;
    free_dict(h);
}

// Column groups of the right pass: consecutive entries of `order` in groups of
// G, the union of each group's columns with their ELL neighbours, and per
// member the LDS slots of itself and its neighbours.  G shrinks until every
// union fits the pass's LDS tile.
int build_right_groups(lmh::DevHam &h, const std::vector<int32_t> &col, const std::vector<int32_t> &order) {
    const int64_t n = h.n;
    const int zs = h.zmax + 1;
    for (int G = lmh::right_pass_group_size(); G >= 1; G /= 2) {
        const int64_t ng = (n + G - 1) / G;
        std::vector<int32_t> ucols;
        std::vector<int64_t> uptr(ng + 1, 0);
        std::vector<int16_t> umap((size_t)n * zs, -1);
        int64_t umax = 0;
        bool fits = true;
        std::vector<int32_t> u;
        for (int64_t g = 0; g < ng && fits; ++g) {
            const int64_t m0 = g * G, m1 = std::min<int64_t>(m0 + G, n);
            u.clear();
            for (int64_t m = m0; m < m1; ++m) {
                const int32_t j = order[m];
                u.push_back(j);
                for (int q = 0; q < h.zmax; ++q) {
                    const int32_t l = col[(int64_t)q * n + j];
                    if (l >= 0) u.push_back(l);
                }
            }
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            if ((int64_t)u.size() > lmh::right_pass_max_union()) {
                fits = false;
                break;
            }
            umax = std::max<int64_t>(umax, (int64_t)u.size());
            auto slot = [&](int32_t c) { return (int16_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin()); };
            for (int64_t m = m0; m < m1; ++m) {
                const int32_t j = order[m];
                umap[(size_t)m * zs] = slot(j);
                for (int q = 0; q < h.zmax; ++q) {
                    const int32_t l = col[(int64_t)q * n + j];
                    umap[(size_t)m * zs + 1 + q] = l >= 0 ? slot(l) : (int16_t)-1;
                }
            }
            uptr[g] = (int64_t)ucols.size();
            ucols.insert(ucols.end(), u.begin(), u.end());
        }
        if (!fits) continue;
        uptr[ng] = (int64_t)ucols.size();
        dfree(h.gcols);
        dfree(h.ucols);
        dfree(h.uptr);
        dfree(h.umap);
        hipError_t e = dalloc(&h.gcols, n);
        if (e == hipSuccess) e = dalloc(&h.ucols, ucols.size());
        if (e == hipSuccess) e = dalloc(&h.uptr, ng + 1);
        if (e == hipSuccess) e = dalloc(&h.umap, umap.size());
        if (e == hipSuccess) e = hipMemcpy(h.gcols, order.data(), 4 * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.ucols, ucols.data(), 4 * ucols.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.uptr, uptr.data(), 8 * (ng + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.umap, umap.data(), 2 * umap.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = lmh::prepare_passes(nullptr);
        if (e != hipSuccess) return hipfail(e, "right-pass groups");
        h.ngroups = ng;
        h.umax = umax;
        h.group = G;
        return LMH_OK;
    }
    return fail(LMH_ERANGE, "right-pass column groups do not fit the LDS tile");
}

// Z-order (Morton) permutation of sites by their coordinates rounded to
// lattice units; ties keep index order.
std::vector<int32_t> zorder(const double *xy, int64_t n) {
    double xmin = xy[0], ymin = xy[1];
    for (int64_t i = 0; i < n; ++i) {
        xmin = std::min(xmin, xy[2 * i]);
        ymin = std::min(ymin, xy[2 * i + 1]);
    }
    auto spread = [](uint64_t v) {
        v &= 0xffffffffull;
        v = (v | (v << 16)) & 0x0000ffff0000ffffull;
        v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
        v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int32_t>> keys(n);
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t qx = (uint64_t)std::llround(xy[2 * i] - xmin);
        const uint64_t qy = (uint64_t)std::llround(xy[2 * i + 1] - ymin);
        keys[i] = {spread(qx) | (spread(qy) << 1), (int32_t)i};
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = keys[i].second;
    return order;
}

}  // namespace lmhapi

extern "C" {
int lmh_set_geometry(lmh_context *ctx, int64_t nsites, int64_t nbonds, const int64_t *bond_i,
                     const int64_t *bond_j, const double *r_i, const double *r_j, const double *amp,
                     const double *twist) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nsites <= 0 || nsites >= (int64_t(1) << 31)) return fail(LMH_EINVAL, "invalid number of sites %lld", (long long)nsites);
    if (nbonds < 0) return fail(LMH_EINVAL, "negative number of bonds");
    if (nbonds > 0 && (!bond_i || !bond_j || !r_i || !r_j || !amp))
        return fail(LMH_EINVAL, "null bond array");
    for (lmh_state *st : ctx->states)
        if (st->n != nsites) return fail(LMH_ESTATE, "existing states have %lld sites", (long long)st->n);
    for (int64_t b = 0; b < nbonds; ++b) {
        if (bond_i[b] < 0 || bond_i[b] >= nsites || bond_j[b] < 0 || bond_j[b] >= nsites)
            return fail(LMH_EINVAL, "bond %lld has a site index out of range", (long long)b);
        if (bond_i[b] == bond_j[b])
            return fail(LMH_EINVAL, "bond %lld connects site %lld to itself", (long long)b, (long long)bond_i[b]);
    }
    // deferred passes run with the geometry they were made for
    rc = ctx_settle_all(ctx);
    if (rc) return rc;
    // ELL slots: distinct columns of each row, sorted; contributions in bond order
    const int64_t n = nsites;
    std::vector<std::vector<int64_t>> rowcols(n);
    for (int64_t b = 0; b < nbonds; ++b) {
        rowcols[bond_i[b]].push_back(bond_j[b]);
        rowcols[bond_j[b]].push_back(bond_i[b]);
    }
    int zmax = 0;
    for (auto &rc_ : rowcols) {
        std::sort(rc_.begin(), rc_.end());
        rc_.erase(std::unique(rc_.begin(), rc_.end()), rc_.end());
        zmax = std::max<int>(zmax, (int)rc_.size());
    }
    auto slot_of = [&](int64_t r, int64_t c) -> int64_t {
        const auto &v = rowcols[r];
        int64_t s = std::lower_bound(v.begin(), v.end(), c) - v.begin();
        return s * n + r;
    };
    const int64_t nslots = (int64_t)zmax * n;
    std::vector<int32_t> col(nslots, -1);
    for (int64_t r = 0; r < n; ++r)
        for (size_t s = 0; s < rowcols[r].size(); ++s) col[(int64_t)s * n + r] = (int32_t)rowcols[r][s];
    int64_t band = 0;
    for (int64_t q = 0; q < nslots; ++q)
        if (col[q] >= 0) band = std::max<int64_t>(band, std::llabs((int64_t)col[q] - q % n));
    std::vector<std::vector<int64_t>> contrib(nslots);
    std::vector<int64_t> bslot(nbonds);
    for (int64_t b = 0; b < nbonds; ++b) {
        const int64_t q = slot_of(bond_i[b], bond_j[b]);
        contrib[q].push_back(b);
        bslot[b] = q;
        contrib[slot_of(bond_j[b], bond_i[b])].push_back(~b);
    }
    std::vector<int64_t> cptr(nslots + 1, 0), cbond;
    cbond.reserve(2 * nbonds);
    for (int64_t q = 0; q < nslots; ++q) {
        cptr[q] = (int64_t)cbond.size();
        cbond.insert(cbond.end(), contrib[q].begin(), contrib[q].end());
    }
    cptr[nslots] = (int64_t)cbond.size();
    std::vector<int32_t> bi(nbonds), bj(nbonds);
    std::vector<double> tw(2 * nbonds);
    for (int64_t b = 0; b < nbonds; ++b) {
        bi[b] = (int32_t)bond_i[b];
        bj[b] = (int32_t)bond_j[b];
        tw[2 * b] = twist ? twist[2 * b] : 1.0;
        tw[2 * b + 1] = twist ? twist[2 * b + 1] : 0.0;
    }
    // (re)allocate
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    csc_release(ctx);
    std::vector<double> onsite_keep = ctx->onsite_host;
    tape_clear(ctx);
    pool_free(ctx);
    free_geometry(ctx->h);
    lmh::DevHam &h = ctx->h;
    h.n = n;
    h.nbonds = nbonds;
    h.zmax = zmax;
    h.band = band;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = dalloc(&h.ri, nbonds);
    if (e == hipSuccess) e = dalloc(&h.rj, nbonds);
    if (e == hipSuccess) e = dalloc(&h.amp, nbonds);
    if (e == hipSuccess) e = dalloc(&h.twist, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bslot, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bi, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bj, nbonds);
    if (e == hipSuccess) e = dalloc(&h.cptr, nslots + 1);
    if (e == hipSuccess) e = dalloc(&h.cbond, cbond.size());
    const int64_t zq = (zmax + 3) / 4;
    std::vector<int32_t> col4(zq * n * 4, -1);
    for (int q = 0; q < zmax; ++q)
        for (int64_t i = 0; i < n; ++i) col4[((q / 4) * n + i) * 4 + q % 4] = col[q * n + i];
    if (e == hipSuccess) e = dalloc(&h.col, nslots);
    if (e == hipSuccess) e = dalloc(&h.col4, col4.size());
    // val padded to whole 4-slot chunks of col4; the padding stays zero
    if (e == hipSuccess) e = dalloc(&h.val, zq * 4 * n);
    if (e == hipSuccess) e = hipMemset(h.val, 0, sizeof(double2) * zq * 4 * n);
    if (e == hipSuccess) e = dalloc(&h.diag, n);
    if (e == hipSuccess && zmax <= 4 && n < 65536) e = dalloc(&h.rec, n);
    if (e != hipSuccess) {
        free_geometry(ctx->h);
        ctx->geom_set = false;
        return hipfail(e, "geometry allocation");
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (bytes && e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    up(h.ri, r_i, 16 * nbonds);
    up(h.rj, r_j, 16 * nbonds);
    up(h.amp, amp, 16 * nbonds);
    up(h.twist, tw.data(), 16 * nbonds);
    up(h.bslot, bslot.data(), 8 * nbonds);
    up(h.bi, bi.data(), 4 * nbonds);
    up(h.bj, bj.data(), 4 * nbonds);
    up(h.cptr, cptr.data(), 8 * (nslots + 1));
    up(h.cbond, cbond.data(), 8 * cbond.size());
    up(h.col, col.data(), 4 * nslots);
    up(h.col4, col4.data(), 4 * col4.size());
    if (e == hipSuccess) e = lmh::prepare_passes(nullptr);
    // grid-stencil pattern: the strip kernel's geometry and direction slots
    if (e == hipSuccess && n <= 10240) {
        std::vector<int32_t> sslot(4 * n);
        lmh::StripGeom sg;
        if (lmh::strip_detect(col.data(), n, zmax, &sg, sslot.data())) {
            // per site: the direction of each ELL slot (2 bits) and which
            // slots hold a bond (bits 8 + slot), for the grid-window order search
            std::vector<uint32_t> sdir(n, 0);
            for (int d = 0; d < 4; ++d)
                for (int64_t i = 0; i < n; ++i) {
                    const int32_t q = sslot[d * n + i];
                    if (q < 0) continue;
                    const int u = (int)(q / n);
                    sdir[i] |= (uint32_t)d << (2 * u) | 1u << (8 + u);
                }
            e = dalloc(&h.sdir, n);
            if (e == hipSuccess) e = hipMemcpy(h.sdir, sdir.data(), 4 * n, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = dalloc(&h.sslot, sslot.size());
            if (e == hipSuccess) e = dalloc(&h.stval, (size_t)4 * sg.nt);
            if (e == hipSuccess) e = dalloc(&h.sdiag, (size_t)sg.s * sg.nt);
            if (e == hipSuccess) e = hipMemcpy(h.sslot, sslot.data(), 4 * sslot.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                h.sg = sg;
                e = lmh::prepare_strip(h);
            }
        }
    }
    if (e != hipSuccess) {
        free_geometry(ctx->h);
        ctx->geom_set = false;
        return hipfail(e, "geometry upload");
    }
    ctx->geom_set = true;
    ctx->param_version++;
    ctx->geom_version++;
    // schedule options pinned (or autotuned) for the previous geometry that do
    // not fit this one go back to their by-size defaults: they only choose
    // kernels, never results
    if (ctx->fused_variant >= 0 && !lmh::fused_variant_valid(ctx->fused_variant, h.n, h.zmax)) ctx->fused_variant = -1;
    if ((ctx->left_variant == 2 && !lmh::left_band_fits(h)) || (ctx->left_variant == 3 && !lmh::left_band2_fits(h)))
        ctx->left_variant = 0;
    ctx->solver_ready = false;
    dfree(ctx->d_rho);
    dfree(ctx->d_J);
    dfree(ctx->d_hdiag);
    dfree(ctx->d_hij);
    if (ctx->onsite_set && (int64_t)onsite_keep.size() == n) {
        HIP_TRY(dalloc(&h.onsite, n), "on-site allocation");
        HIP_TRY(hipMemcpy(h.onsite, onsite_keep.data(), 8 * n, hipMemcpyHostToDevice), "on-site upload");
    } else {
        ctx->onsite_set = false;
        ctx->onsite_host.clear();
    }
    ctx->col_host = col;
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    return build_right_groups(h, ctx->col_host, order);
}

int lmh_set_site_coords(lmh_context *ctx, const double *xy) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_set_site_coords before lmh_set_geometry");
    if (!xy) return fail(LMH_EINVAL, "null coordinates");
    for (int64_t i = 0; i < 2 * ctx->h.n; ++i)
        if (!std::isfinite(xy[i])) return fail(LMH_EINVAL, "coordinate %lld is not finite", (long long)i);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return build_right_groups(ctx->h, ctx->col_host, zorder(xy, ctx->h.n));
}

int lmh_set_onsite(lmh_context *ctx, const double *diag) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_set_onsite before lmh_set_geometry");
    if (ctx->csc_mode) return fail(LMH_ESTATE, "H is given as CSC values: its diagonal comes with lmh_set_csc_values");
    const int64_t n = ctx->h.n;
    if (!diag) {
        if (ctx->onsite_set) {
            HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            dfree(ctx->h.onsite);
            ctx->onsite_set = false;
            ctx->onsite_host.clear();
            ctx->param_version++;
            ctx->onsite_version++;
        }
        return LMH_OK;
    }
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(diag[i])) return fail(LMH_EINVAL, "on-site energy %lld is not finite", (long long)i);
    if (ctx->onsite_set && std::memcmp(ctx->onsite_host.data(), diag, 8 * n) == 0) return LMH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (!ctx->h.onsite) HIP_TRY(dalloc(&ctx->h.onsite, n), "on-site allocation");
    ctx->onsite_host.assign(diag, diag + n);
    HIP_TRY(hipMemcpy(ctx->h.onsite, diag, 8 * n, hipMemcpyHostToDevice), "on-site upload");
    ctx->onsite_set = true;
    ctx->param_version++;
    ctx->onsite_version++;
    return LMH_OK;
}

int lmh_set_field(lmh_context *ctx, int64_t ncomp, const int32_t *kind, const double *params) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ncomp < 0) return fail(LMH_EINVAL, "negative number of field components");
    if (ctx->csc_mode && ncomp > 0)
        return fail(LMH_ESTATE, "H is given as CSC values: the field's phases are part of them");
    if (ncomp > 0 && (!kind || !params)) return fail(LMH_EINVAL, "null field arrays");
    std::vector<lmh::FieldComp> f(ncomp);
    for (int64_t c = 0; c < ncomp; ++c) {
        if (kind[c] < LMH_FIELD_LANDAU || kind[c] > LMH_FIELD_POINT_SINGULAR)
            return fail(LMH_EINVAL, "invalid field kind %d", kind[c]);
        for (int q = 0; q < 3; ++q)
            if (!std::isfinite(params[3 * c + q])) return fail(LMH_EINVAL, "field parameter not finite");
        f[c].kind = kind[c];
        f[c].pad = 0;
        f[c].a = params[3 * c];
        f[c].px = params[3 * c + 1];
        f[c].py = params[3 * c + 2];
    }
    if (f.size() == ctx->field_host.size() &&
        (f.empty() || std::memcmp(f.data(), ctx->field_host.data(), f.size() * sizeof(lmh::FieldComp)) == 0))
        return LMH_OK;
    if (ncomp > ctx->h.field_cap) {
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        dfree(ctx->h.field);
        HIP_TRY(dalloc(&ctx->h.field, ncomp), "field allocation");
        if (ctx->field_stage) (void)hipHostFree(ctx->field_stage);
        ctx->field_stage = nullptr;
        HIP_TRY(hipHostMalloc((void **)&ctx->field_stage, ncomp * sizeof(lmh::FieldComp), 0), "field staging");
        ctx->h.field_cap = ncomp;
    }
    if (ncomp) {
        // in stream order (the launches already queued read the earlier
        // values), through pinned memory: no wait for the GPU here, beyond
        // the previous upload having left the staging buffer
        if (!ctx->ev_field) HIP_TRY(hipEventCreateWithFlags(&ctx->ev_field, hipEventDisableTiming), "hipEventCreate");
        else HIP_TRY(hipEventSynchronize(ctx->ev_field), "field staging");
        std::memcpy(ctx->field_stage, f.data(), ncomp * sizeof(lmh::FieldComp));
        HIP_TRY(hipMemcpyAsync(ctx->h.field, ctx->field_stage, ncomp * sizeof(lmh::FieldComp), hipMemcpyHostToDevice,
                               ctx->stream),
                "field upload");
        HIP_TRY(hipEventRecord(ctx->ev_field, ctx->stream), "hipEventRecord");
    }
    ctx->h.nfield = ncomp;
    ctx->field_host = f;
    ctx->param_version++;
    return LMH_OK;
}

int lmh_set_csc_pattern(lmh_context *ctx, int64_t n, const int64_t *colptr, const int64_t *rowval, int32_t base) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (base != 0 && base != 1) return fail(LMH_EINVAL, "index base must be 0 or 1");
    if (n <= 0 || n >= (int64_t(1) << 31)) return fail(LMH_EINVAL, "invalid matrix size %lld", (long long)n);
    if (!colptr) return fail(LMH_EINVAL, "null colptr");
    if (colptr[0] != base) return fail(LMH_EINVAL, "colptr[0] = %lld, expected the index base %d", (long long)colptr[0], base);
    for (int64_t c = 0; c < n; ++c)
        if (colptr[c + 1] < colptr[c]) return fail(LMH_EINVAL, "colptr decreases at column %lld", (long long)c);
    const int64_t nnz = colptr[n] - base;
    if (nnz > 0 && !rowval) return fail(LMH_EINVAL, "null rowval");
    for (int64_t c = 0; c < n; ++c)
        for (int64_t k = colptr[c] - base; k < colptr[c + 1] - base; ++k) {
            const int64_t r = rowval[k] - base;
            if (r < 0 || r >= n) return fail(LMH_EINVAL, "row index %lld out of range in column %lld", (long long)rowval[k], (long long)c);
            if (k > colptr[c] - base && r <= rowval[k - 1] - base)
                return fail(LMH_EINVAL, "row indices of column %lld are not strictly increasing", (long long)c);
        }
    // position of entry (r, c) in the value array, -1 when not stored
    auto find = [&](int64_t r, int64_t c) -> int64_t {
        const int64_t *lo = rowval + (colptr[c] - base), *hi = rowval + (colptr[c + 1] - base);
        const int64_t *p = std::lower_bound(lo, hi, r + base);
        return (p != hi && *p == r + base) ? (int64_t)(p - rowval) : -1;
    };
    std::vector<int64_t> bi, bj, kup, klo, kd(n, -1);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t k = colptr[c] - base; k < colptr[c + 1] - base; ++k) {
            const int64_t r = rowval[k] - base;
            if (r == c) {
                kd[c] = k;
                continue;
            }
            const int64_t kt = find(c, r);
            if (kt < 0)
                return fail(LMH_EINVAL, "entry (%lld, %lld) is stored but (%lld, %lld) is not: a Hermitian H has a "
                                        "symmetric pattern", (long long)r, (long long)c, (long long)c, (long long)r);
            if (r < c) {
                bi.push_back(r);
                bj.push_back(c);
                kup.push_back(k);
                klo.push_back(kt);
            }
        }
    const int64_t nb = (int64_t)bi.size();
    std::vector<double> zero(2 * std::max<int64_t>(nb, 1), 0.0);
    rc = lmh_set_geometry(ctx, n, nb, bi.data(), bj.data(), zero.data(), zero.data(), zero.data(), nullptr);
    if (rc) return rc;
    // no field and no parametric on-site terms: the values carry both
    ctx->field_host.clear();
    ctx->h.nfield = 0;
    dfree(ctx->h.onsite);
    ctx->onsite_host.clear();
    hipError_t e = dalloc(&ctx->h.onsite, n);
    if (e == hipSuccess) e = dalloc(&ctx->d_kup, std::max<int64_t>(nb, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_klo, std::max<int64_t>(nb, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_kd, n);
    if (e == hipSuccess) e = dalloc(&ctx->d_nz, std::max<int64_t>(nnz, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_cscflag, 1);
    if (e == hipSuccess && nb) e = hipMemcpy(ctx->d_kup, kup.data(), 8 * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess && nb) e = hipMemcpy(ctx->d_klo, klo.data(), 8 * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->d_kd, kd.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ctx->h.onsite, 0, 8 * n);
    if (e != hipSuccess) {
        csc_release(ctx);
        ctx->onsite_set = false;
        return hipfail(e, "CSC pattern upload");
    }
    ctx->onsite_set = true;
    ctx->csc_mode = true;
    ctx->csc_values = false;
    ctx->csc_nnz = nnz;
    ctx->param_version++;
    ctx->onsite_version++;
    return LMH_OK;
}

int lmh_set_csc_values(lmh_context *ctx, int64_t nnz, const double *nzval) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->csc_mode) return fail(LMH_ESTATE, "lmh_set_csc_values before lmh_set_csc_pattern");
    if (nnz != ctx->csc_nnz)
        return fail(LMH_EINVAL, "%lld values for a pattern of %lld entries", (long long)nnz, (long long)ctx->csc_nnz);
    if (nnz > 0 && !nzval) return fail(LMH_EINVAL, "null values");
    lmh::DevHam &h = ctx->h;
    int flag = 0;
    hipError_t e = nnz ? hipMemcpyAsync(ctx->d_nz, nzval, 16 * nnz, hipMemcpyHostToDevice, ctx->stream) : hipSuccess;
    if (e == hipSuccess)
        e = lmh::launch_csc_gather(h.nbonds, h.n, ctx->d_kup, ctx->d_klo, ctx->d_kd, ctx->d_nz, h.amp, h.onsite,
                                   ctx->d_cscflag, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, ctx->d_cscflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->param_version++;
    ctx->onsite_version++;
    ctx->csc_values = false;
    if (e != hipSuccess) return hipfail(e, "CSC value upload");
    if (flag & 4) return fail(LMH_EINVAL, "H has a value that is not finite");
    if (flag & 1) return fail(LMH_EINVAL, "H is not Hermitian: a stored H[j,i] is not the conjugate of H[i,j]");
    if (flag & 2) return fail(LMH_EINVAL, "H is not Hermitian: a diagonal entry has a nonzero imaginary part");
    ctx->csc_values = true;
    return LMH_OK;
}

}  // extern "C"
namespace lmhapi {

// Takes up an order result: the larger neighbourhood tables when a column's
// did not fit (one level at a time, each read back), then the value-class
// and strip checks that travel with it, K and s.  An error leaves the solver
// not ready (the next lmh_update_solver recomputes).
int order_finish(lmh_context *ctx, lmh::TaylorOut &out, int level0) {
    lmh::DevHam &h = ctx->h;
    int level = level0;
    while (out.overflow && level + 1 < lmh::taylor_order_levels()) {
        ++level;
        {
            ProfScope ps(ctx, LMH_KERNEL_TAYLOR_ORDER);
            HIP_TRY(lmh::launch_taylor_order(h, ctx->dt, ctx->nztol, ctx->d_run, level, ctx->d_partmax,
                                             ctx->d_maxbits, ctx->d_overflow, ctx->stream),
                    "taylor_order launch");
            HIP_TRY(lmh::launch_taylor_finish(h.red, ctx->d_run, ctx->d_maxbits, ctx->d_overflow, ctx->d_tout,
                                              ctx->stream),
                    "taylor finish launch");
        }
        HIP_TRY(hipMemcpyAsync(&out, ctx->d_tout, sizeof out, hipMemcpyDeviceToHost, ctx->stream), "order readback");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    if (!out.overflow) {
        ctx->order_level = level;
        ctx->order_level_geom = ctx->geom_version + 1;
    }
    const unsigned long long *red = out.red;
    h.dict_ok = h.ndict > 0 && red[2] == 0;
    h.strip_ok = h.sg.s > 0 && (red[3] & 1) == 0;
    h.strip_xreal = h.strip_ok && (red[3] & 2) == 0;
    ctx->flags_geom = ctx->geom_version + 1;
    if (out.overflow) {
        ctx->solver_ready = false;
        return fail(LMH_ERANGE, "Taylor-order search: neighbourhood of %d terms exceeds the kernel capacity",
                    out.nmax);
    }
    if (out.K < 0) {
        ctx->solver_ready = false;
        return fail(LMH_ERANGE, "Taylor series did not reach the threshold within %d terms", out.nmax);
    }
    ctx->K = out.K;
    ctx->s = out.s;
    return LMH_OK;
}

// the step that follows runs on the strip kernel with a deferred second pass
// (the path that can take its polynomial from the device), by the flags of
// the last result taken up for this geometry
bool strip_step_path(const lmh_context *ctx) {
    const lmh::DevHam &h = ctx->h;
    int mode = ctx->step_mode;
    if (mode == LMH_STEP_AUTO)
        mode = (lmh::fused_supported(h) && h.n <= ctx->fused_max_n) ? LMH_STEP_FUSED_T : LMH_STEP_PER_PASS;
    return mode == LMH_STEP_FUSED_T && lmh::fused_supported(h) && ctx->defer && h.strip_ok &&
           lmh::fused_resolved_variant(h, ctx->fused_variant) == lmh::fused_strip_variant();
}

bool order_may_stay(const lmh_context *ctx) {
    return ctx->async_order && ctx->flags_geom == ctx->geom_version + 1 && !ctx->tape.active && strip_step_path(ctx);
}

}  // namespace lmhapi
extern "C" {

int lmh_update_solver(lmh_context *ctx, double dt, int force, double threshold, double nztol, int32_t *K_out,
                      double *s_out) {
    // (an earlier update's result still on the device is taken up here: by
    // now the launches in front of it have normally finished)
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_update_solver before lmh_set_geometry");
    if (ctx->csc_mode && !ctx->csc_values)
        return fail(LMH_ESTATE, "lmh_update_solver without valid lmh_set_csc_values for the CSC pattern");
    if (!std::isfinite(dt) || !(threshold > 0) || !(nztol >= 0))
        return fail(LMH_EINVAL, "invalid dt / threshold / nztol");
    const bool same = ctx->solver_ready && !force && ctx->solver_version == ctx->param_version &&
                      isapprox(-dt, -ctx->dt) && threshold == ctx->threshold && nztol == ctx->nztol;
    if (!same) {
        lmh::DevHam &h = ctx->h;
        if (ctx->dict_geom != ctx->geom_version || ctx->dict_probes_built != ctx->dict_probes ||
            ctx->dict_sig != dict_signature(ctx))
            HIP_TRY(build_dict(ctx), "value classes");
        if (ctx->built_version != ctx->param_version) {
            ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
            HIP_TRY(lmh::launch_build_h(h, ctx->stream), "build_h launch");
            ctx->built_version = ctx->param_version;
        }
        // the order rule on the device: s and the order bound from H's
        // reductions, the per-column term maxima, K and the Horner
        // coefficients
        const int level0 = ctx->order_level_geom == ctx->geom_version + 1 ? ctx->order_level : 0;
        {
            ProfScope ps(ctx, LMH_KERNEL_TAYLOR_ORDER);
            HIP_TRY(lmh::launch_taylor_setup(h.red, dt, threshold, ctx->d_run, ctx->d_overflow, ctx->stream),
                    "taylor setup launch");
            HIP_TRY(lmh::launch_taylor_order(h, dt, nztol, ctx->d_run, level0, ctx->d_partmax, ctx->d_maxbits,
                                             ctx->d_overflow, ctx->stream),
                    "taylor_order launch");
            HIP_TRY(lmh::launch_taylor_finish(h.red, ctx->d_run, ctx->d_maxbits, ctx->d_overflow, ctx->d_tout,
                                              ctx->stream),
                    "taylor finish launch");
        }
        ctx->dt = dt;
        ctx->threshold = threshold;
        ctx->nztol = nztol;
        if (!K_out && !s_out && order_may_stay(ctx)) {
            // K stays on the device: the strip steps read it from there; the
            // result comes back behind an event and is taken up later
            if (!ctx->h_tout) HIP_TRY(hipHostMalloc((void **)&ctx->h_tout, sizeof(lmh::TaylorOut), 0), "pinned result");
            if (!ctx->ev_order)
                HIP_TRY(hipEventCreateWithFlags(&ctx->ev_order, hipEventDisableTiming), "hipEventCreate");
            HIP_TRY(hipMemcpyAsync(ctx->h_tout, ctx->d_tout, sizeof(lmh::TaylorOut), hipMemcpyDeviceToHost,
                                   ctx->stream),
                    "order readback");
            HIP_TRY(hipEventRecord(ctx->ev_order, ctx->stream), "hipEventRecord");
            ctx->order_async = true;
            ctx->async_xreal = h.strip_xreal ? 1 : 0;
            ctx->async_level = level0;
            ctx->solver_version = ctx->param_version;
            ctx->solver_serial++;
            ctx->solver_ready = true;
            return LMH_OK;
        }
        // one readback of the result (with the reductions and the value-class
        // / strip checks)
        lmh::TaylorOut out;
        HIP_TRY(hipMemcpyAsync(&out, ctx->d_tout, sizeof out, hipMemcpyDeviceToHost, ctx->stream), "order readback");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        rc = order_finish(ctx, out, level0);
        if (rc) return rc;
        ctx->solver_version = ctx->param_version;
        ctx->solver_serial++;
        ctx->solver_ready = true;
    }
    if (K_out) *K_out = ctx->K;
    if (s_out) *s_out = ctx->s;
    return LMH_OK;
}

int lmh_solver_order(lmh_context *ctx, int32_t *K, double *s) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->solver_ready) return fail(LMH_ESTATE, "lmh_solver_order before lmh_update_solver");
    if (K) *K = ctx->K;
    if (s) *s = ctx->s;
    return LMH_OK;
}

}  // extern "C"
