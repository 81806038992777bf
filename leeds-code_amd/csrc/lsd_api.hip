// C-ABI entry points (include/lsd.h): context life cycle, field I/O and the
// phases of the reference's main loop (program/main.F90:100-164).
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lsd_internal.h"

namespace lsd {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static int check_config(const lsd_config* c, int* H, int* Th, int* Ph) {
    if (!c) {
        set_error("null config");
        return LSD_E_ARG;
    }
    // Scripts/parameters.F90_template:274-288
    if (c->L <= (c->M - 1) * c->Mp) { set_error("set L > (M-1)*Mp"); return LSD_E_CONFIG; }
    if (c->M > 1 && (c->M % 2) == 1) { set_error("if(M>1) set M even"); return LSD_E_CONFIG; }
    if (c->Mp < 1 || c->M < 1) { set_error("M, Mp must be positive"); return LSD_E_CONFIG; }
    if (c->N < 7 || c->N > 1024) { set_error("N must be in [7, 1024]"); return LSD_E_CONFIG; }
    int th = c->Th > 0 ? c->Th : c->L * 3 / 2;
    int ph = c->Ph > 0 ? c->Ph : (c->M * 3 / 2) * 2;
    if (th % 2) { set_error("Th must be even"); return LSD_E_CONFIG; }
    if (ph < 2 * c->M || ph % 2) { set_error("Ph must be even and >= 2M"); return LSD_E_CONFIG; }
    if (!c->vel) { set_error("vel must be time-stepped"); return LSD_E_CONFIG; }
    if (c->vel_bc < 1 || c->vel_bc > 4 || c->cod_bc != 1 || c->comp_bc != 1) {
        set_error("i_vel_bc must be 1..4, i_cod_bc = i_comp_bc = 1");
        return LSD_E_CONFIG;
    }
    if (c->Nic > 1 && c->vel_bc > 2) {
        set_error("a conducting inner core needs a no-slip inner boundary (i_vel_bc = 1 or 2)");
        return LSD_E_CONFIG;
    }
    if (c->dt == 0.0) { set_error("dt must be nonzero (> 0 fixed, < 0 adaptive)"); return LSD_E_CONFIG; }
    if (c->step_control < 0 || c->step_control > 2) {
        set_error("i_time_step_control must be 1 or 2 (0 selects 2)");
        return LSD_E_CONFIG;
    }
    if (c->precision < LSD_PLAIN_F64 || c->precision > LSD_MIXED_F32_MAG64) {
        set_error("precision must be LSD_PLAIN_F64, LSD_MIXED_F32, LSD_MIXED_BF16X3 or LSD_MIXED_F32_MAG64");
        return LSD_E_CONFIG;
    }
    if (c->precision != LSD_PLAIN_F64 && ph > 1024) {
        set_error("the mixed forms need Ph <= 1024 (fused physical-space pass)");
        return LSD_E_CONFIG;
    }
    if (c->Nic < 0 || (c->Nic > 1 && (!c->mag || c->Nic < 8 || c->Nic > c->N))) {
        set_error("i_Nic > 1 needs the magnetic field and 8 <= i_Nic <= i_N");
        return LSD_E_CONFIG;
    }
    if (c->rot != 0 && c->rot != 1) {
        set_error("rot (b_rot_tstep) must be 0 or 1");
        return LSD_E_CONFIG;
    }
    if (H) *H = c->L * c->M - c->Mp * (c->M - 1) * c->M / 2;
    if (Th) *Th = th;
    if (Ph) *Ph = ph;
    return LSD_OK;
}

// Folded harmonic order (Scripts/parallel.h_template:65-80).
void harmonics(int L, int M, int Mp, std::vector<int>& hl, std::vector<int>& hm) {
    int H = L * M - Mp * (M - 1) * M / 2;
    int Hc = (M + 1) / 2, Hr = H / Hc;
    hl.resize(H);
    hm.resize(H);
    int nh = 0;
    for (int jj = 1; jj <= Hc; jj++)
        for (int ii = 1; ii <= Hr; ii++) {
            int m = jj - 1, l = m * Mp + ii - 1;
            if (l >= L) {
                m = 2 * Hc - jj;
                l = L - Hr + ii - 1;
            }
            hl[nh] = l;
            hm[nh] = m * Mp;
            nh++;
        }
}

template <typename T>
static int dalloc(T** p, size_t n) {
    LSD_HIP(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
    return LSD_OK;
}

static void drop_graphs(Ctx& c);

static void release(Ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    void* ptrs[] = {c->d_hl, c->d_hm, c->d_pidx, c->d_apos, c->d_slots, c->d_slot_order, c->d_mslot, c->d_gp, c->d_gw,
                    c->d_sin, c->d_tab, c->d_atab, c->d_r, c->d_dr1, c->d_dr2, c->d_radlap, c->d_r1dr,
                    c->d_intr2dr, c->d_ll, c->d_lu, c->d_piv, c->d_pivpk, c->d_y, c->d_gre, c->d_invg, c->d_info,
                    c->d_lut32, c->d_xband, c->d_sblk[0], c->d_sblk[1], c->d_sblk[2], c->d_sord[0], c->d_sord[1], c->d_sord[2], c->d_tab32,
                    c->d_atab32, c->d_tw32, c->d_slot_t0, c->d_row_h, c->d_atab4, c->d_ana_base, c->d_acol,
                    c->d_spec_in, c->d_spec_out, c->d_fourier, c->d_folded, c->d_phys, c->d_nlphys,
                    c->d_tw, c->d_partial, c->d_cfl_d, c->d_cfl_part, c->d_cfl, c->d_tabb, c->d_atabb,
                    c->d_bx_sbase, c->d_icr, c->d_icdr1, c->d_icdr2, c->d_icradlap, c->d_iclu, c->d_icpiv,
                    c->d_icy, c->d_icblk, c->d_icord, c->d_icpartial, c->d_rot, c->d_gwsin,
                    c->d_icf[0], c->d_icf[1], c->d_icn[0], c->d_icn[1], c->d_ichist[0], c->d_ichist[1],
                    c->d_wq};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int k = 0; k < LSD_NUM_COLL; k++)
        if (c->d_f[k]) (void)hipFree(c->d_f[k]);
    for (int k = 0; k < 6; k++) {
        if (c->d_hist[k]) (void)hipFree(c->d_hist[k]);
        if (c->d_rhs[k]) (void)hipFree(c->d_rhs[k]);
    }
    if (c->d_rhs2) (void)hipFree(c->d_rhs2);
    if (c->h_partial) (void)hipHostFree(c->h_partial);
    if (c->h_icpartial) (void)hipHostFree(c->h_icpartial);
    if (c->h_cfl) (void)hipHostFree(c->h_cfl);
    for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
    drop_graphs(*c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// ------------------------------------------------------- tuning cache
//
// One JSON object per line, written by lsd_tune:
//   {"L": 170, "M": 170, "Mp": 1, "N": 96, "Th": 256, "Ph": 512,
//    "precision": 1, "syn_scalar": 4, "syn_vector": 8, "syn_grad": 8,
//    "ana": 0, "solve": 0}
// The last line matching a context's sizes and precision applies.

static std::string tuning_path() {
    if (const char* e = std::getenv("LSD_TUNING_CACHE")) return e;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&tuning_path), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t k = p.rfind('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/tuning_cache.jsonl";
    }
    return "tuning_cache.jsonl";
}

struct TuneEntry {
    int L, M, Mp, N, Th, Ph, prec, v[LSD_V_COUNT];
};

static bool parse_tune_line(const char* line, TuneEntry& e) {
    return std::sscanf(line,
                       " {\"L\": %d, \"M\": %d, \"Mp\": %d, \"N\": %d, \"Th\": %d, \"Ph\": %d, "
                       "\"precision\": %d, \"syn_scalar\": %d, \"syn_vector\": %d, \"syn_grad\": %d, "
                       "\"ana\": %d, \"solve\": %d",
                       &e.L, &e.M, &e.Mp, &e.N, &e.Th, &e.Ph, &e.prec, &e.v[0], &e.v[1], &e.v[2], &e.v[3],
                       &e.v[4]) == 12;
}

static void select_solve_form(Ctx& c);

// Whether a variant value applies to this context.
static bool variant_ok(const Ctx& c, int which, int value) {
    switch (which) {
        case LSD_V_SYN_SCALAR:
        case LSD_V_SYN_VECTOR:
        case LSD_V_SYN_GRAD: return (value >= 4 && value <= 6) || (value >= 8 && value <= 10);
        case LSD_V_ANA: return value == 1 || ((value == 0 || value == 2) && c.ana_wide_ok);
        case LSD_V_SOLVE:
            return (value == 0 && c.solve_direct_ok) || (value == 1 && c.hb_form[1] >= 1) ||
                   (value == 2 && c.mixed && c.hb_form[2] >= 1);
        default: return false;
    }
}

static void load_tuning(Ctx& c) {
    if (!c.ana_wide_ok) c.var[LSD_V_ANA] = 1;
    FILE* f = std::fopen(tuning_path().c_str(), "r");
    if (!f) return;
    char line[1024];
    TuneEntry e, hit;
    bool found = false, own = false;
    while (std::fgets(line, sizeof line, f)) {
        if (!(parse_tune_line(line, e) && e.L == c.L && e.M == c.M && e.Mp == c.Mp && e.N == c.N && e.Th == c.Th &&
              e.Ph == c.Ph))
            continue;
        if (e.prec == c.cfg.precision) {
            hit = e;
            found = own = true;
        } else if (c.cfg.precision == LSD_MIXED_F32_MAG64 && e.prec == LSD_MIXED_F32 && !own) {
            // the FP64-magnetic form runs the FP32 form's kernels for everything
            // else, so it takes that form's choices when it has none of its own
            hit = e;
            found = true;
        }
    }
    std::fclose(f);
    if (!found) return;
    for (int k = 0; k < LSD_V_COUNT; k++)
        if (variant_ok(c, k, hit.v[k])) c.var[k] = hit.v[k];
    select_solve_form(c);
}

// Points the solve kernels at the block set of the selected form.
static void select_solve_form(Ctx& c) {
    const int v = c.var[LSD_V_SOLVE];
    const int form = (v == 0 && c.solve_direct_ok) ? 0 : (v == 2 && c.hb_form[2] >= 1) ? 2 : 1;
    c.d_sblocks = c.d_sblk[form];
    c.d_sorder = c.d_sord[form];
    c.n_partial_blocks = c.nblk_form[form];
    c.solve_hb = c.hb_form[form];
}

static int create(const lsd_config* cfg, int device, Ctx** out) {
    int H, Th, Ph;
    int rc = check_config(cfg, &H, &Th, &Ph);
    if (rc) return rc;
    Ctx* c = new Ctx();
    c->cfg = *cfg;
    c->device = device;
    c->L = cfg->L;
    c->M = cfg->M;
    c->Mp = cfg->Mp;
    c->N = cfg->N;
    c->Th = Th;
    c->Ph = Ph;
    c->H = H;
    c->Th2 = (Th + 1) / 2;
    c->MP = (c->M + 15) / 16 * 16;
    // d_timestep < 0: adaptive step, starting from |d_timestep| (the dt
    // io_load_state would read from the state file, program/io.F90:429)
    c->adaptive = cfg->dt < 0.0;
    c->dt = std::fabs(cfg->dt);
    c->mixed = cfg->precision != LSD_PLAIN_F64;
    c->mag64 = cfg->precision == LSD_MIXED_F32_MAG64 && cfg->mag;
    c->Nic = cfg->Nic > 1 ? cfg->Nic : 1;
    c->NT = c->Nic + cfg->N - 1;
    *out = c;
    LSD_HIP(hipSetDevice(device));
    LSD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    rc = energy_noalloc_init(device);
    if (rc) return rc;

    // host index bookkeeping
    harmonics(c->L, c->M, c->Mp, c->hl, c->hm);
    for (int h = 0; h < H; h++) {
        if (h == 0 || c->hm[h] != c->hm[h - 1]) c->slots.push_back(Slot{h, 0, c->hm[h], c->hm[h] / c->Mp});
        c->slots.back().len++;
    }
    c->nslot = (int)c->slots.size();
    for (int h = 0; h < H; h++) {
        if (c->hl[h] == 1 && c->hm[h] == 0) c->h10 = h;
        if (c->hl[h] == 1 && c->hm[h] == 1) c->h11 = h;
    }
    c->slot_order.resize(c->nslot);
    for (int s = 0; s < c->nslot; s++) c->slot_order[s] = s;
    std::stable_sort(c->slot_order.begin(), c->slot_order.end(),
                     [&](int a, int b) { return c->slots[a].len > c->slots[b].len; });
    int L = c->L, N = c->N;
    std::vector<int> pidx((size_t)L * (L + 1) / 2, -1), apos(H);
    for (int h = 0; h < H; h++) pidx[(size_t)c->hl[h] * (c->hl[h] + 1) / 2 + c->hm[h]] = h;
    for (auto& s : c->slots)
        for (int k = 0; k < s.len; k++)
            apos[s.h0 + k] = s.h0 + ((k & 1) ? (s.len + 1) / 2 + (k >> 1) : (k >> 1));

    size_t HN = c->coll_elems();
    if (dalloc(&c->d_hl, H) || dalloc(&c->d_hm, H) || dalloc(&c->d_pidx, pidx.size()) ||
        dalloc(&c->d_apos, H) || dalloc(&c->d_slots, c->nslot) || dalloc(&c->d_slot_order, c->nslot) ||
        dalloc(&c->d_mslot, c->M) ||
        dalloc(&c->d_gp, Th) || dalloc(&c->d_gw, Th) || dalloc(&c->d_sin, Th) ||
        dalloc(&c->d_tab, (size_t)5 * H * c->Th2) || dalloc(&c->d_atab, (size_t)3 * H * c->Th2) ||
        dalloc(&c->d_r, (size_t)7 * N) || dalloc(&c->d_dr1, (size_t)7 * N) || dalloc(&c->d_dr2, (size_t)7 * N) ||
        dalloc(&c->d_radlap, (size_t)7 * N) || dalloc(&c->d_r1dr, (size_t)7 * N) || dalloc(&c->d_intr2dr, N) ||
        dalloc(&c->d_ll, (size_t)4 * L) || dalloc(&c->d_lu, (size_t)NLU * L * N * LSD_LDAB) ||
        dalloc(&c->d_piv, (size_t)NLU * L * N) || dalloc(&c->d_y, (size_t)NY * L * N * 7) ||
        dalloc(&c->d_gre, (size_t)2 * L * N) || dalloc(&c->d_invg, (size_t)4 * L) ||
        dalloc(&c->d_info, (size_t)NLU * L) || dalloc(&c->d_spec_in, (size_t)SI_COUNT * HN) ||
        dalloc(&c->d_spec_out, (size_t)SO_COUNT * HN) ||
        dalloc(&c->d_lut32, (size_t)NLU * L * N * LSD_LDAB) || dalloc(&c->d_xband, (size_t)NLU * L * N * 7) ||
        dalloc(&c->d_pivpk, (size_t)NLU * L * ((N + 15) / 16)) ||
        dalloc(&c->d_sorder, H) ||
        dalloc(&c->d_fourier, (size_t)LSD_NUM_PHYS * c->fourier_elems()) ||
        dalloc(&c->d_folded, (size_t)NO_COUNT * c->folded_elems()) ||
        dalloc(&c->d_phys, (size_t)LSD_NUM_PHYS * c->phys_elems()) ||
        dalloc(&c->d_nlphys, (size_t)NO_COUNT * c->phys_elems()))
        return LSD_E_HIP;
    for (int k = 0; k < LSD_NUM_COLL; k++)
        if (dalloc(&c->d_f[k], HN)) return LSD_E_HIP;
    for (int k = 0; k < 6; k++)
        if (dalloc(&c->d_hist[k], HN)) return LSD_E_HIP;
    const bool direct = solve_direct_supported(N);
    if (direct) {
        for (int k = 0; k < 6; k++)
            if (dalloc(&c->d_rhs[k], HN)) return LSD_E_HIP;
        if (dalloc(&c->d_rhs2, HN)) return LSD_E_HIP;
    }
    // solve blocks of both forms: harmonics grouped by degree, at most 32
    // per block for the two-pass kernels, at most what the LDS-column
    // kernels' LDS holds for those
    c->solve_direct_ok = direct;
    c->hb_form[0] = 32;
    c->hb_form[1] = solve_block_harmonics(N, c->mixed);
    c->hb_form[2] = c->mixed ? solve_block_harmonics(N, false) : 0;
    if (c->hb_form[1] < 1 && !direct) {
        set_error("N too large for the banded-solve kernels");
        return LSD_E_CONFIG;
    }
    std::vector<int> sorder[3];
    std::vector<int4> sblocks[3];
    {
        std::vector<std::vector<int>> byl(L);
        for (int h = 0; h < H; h++) byl[c->hl[h]].push_back(h);
        for (int form = 0; form < 3; form++) {
            const int hb = c->hb_form[form];
            if (hb < 1 || (form == 0 && !direct)) continue;
            for (int l = 0; l < L; l++)
                for (size_t k = 0; k < byl[l].size(); k += hb) {
                    int cnt = (int)std::min(byl[l].size() - k, (size_t)hb);
                    sblocks[form].push_back(make_int4(l, (int)sorder[form].size(), cnt, 0));
                    for (int q = 0; q < cnt; q++) sorder[form].push_back(byl[l][k + q]);
                }
            if (dalloc(&c->d_sblk[form], sblocks[form].size()) || dalloc(&c->d_sord[form], H)) return LSD_E_HIP;
            c->nblk_form[form] = (int)sblocks[form].size();
        }
    }
    // without the two-pass kernels the mixed forms solve with FP64 factors
    // by default: the FP32 factors' refinement sweep costs more than it saves
    c->var[LSD_V_SOLVE] = direct ? 0 : (c->hb_form[2] >= 1 ? 2 : 1);
    select_solve_form(*c);
    const int nblk_max = std::max(std::max(c->nblk_form[0], c->nblk_form[1]), c->nblk_form[2]);
    if (dalloc(&c->d_partial, (size_t)8 * nblk_max)) return LSD_E_HIP;
    LSD_HIP(hipHostMalloc((void**)&c->h_partial, sizeof(double) * 8 * nblk_max, hipHostMallocDefault));
    // rotation state; with a conducting inner core its mesh, operators,
    // fields and the blocks of its coupled solves
    if (dalloc(&c->d_rot, ROT_COUNT) || dalloc(&c->d_gwsin, Th) || dalloc(&c->d_wq, 32)) return LSD_E_HIP;
    LSD_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    std::vector<int> icorder;
    std::vector<int4> icblocks;
    if (c->Nic > 1) {
        const int Ni = c->Nic, NT = c->NT;
        c->ic_hb = ic_block_harmonics(N, Ni);
        if (c->ic_hb < 1) {
            set_error("i_N + i_Nic too large for the inner-core solves");
            return LSD_E_CONFIG;
        }
        std::vector<std::vector<int>> byl(L);
        for (int h = 0; h < H; h++) byl[c->hl[h]].push_back(h);
        for (int l = 0; l < L; l++)
            for (size_t k = 0; k < byl[l].size(); k += c->ic_hb) {
                int cnt = (int)std::min(byl[l].size() - k, (size_t)c->ic_hb);
                icblocks.push_back(make_int4(l, (int)icorder.size(), cnt, 0));
                for (int q = 0; q < cnt; q++) icorder.push_back(byl[l][k + q]);
            }
        c->n_icblk = (int)icblocks.size();
        if (dalloc(&c->d_icr, (size_t)7 * Ni) || dalloc(&c->d_icdr1, (size_t)7 * Ni) ||
            dalloc(&c->d_icdr2, (size_t)7 * Ni) || dalloc(&c->d_icradlap, (size_t)7 * Ni) ||
            dalloc(&c->d_iclu, (size_t)2 * L * NT * LSD_LDAB) || dalloc(&c->d_icpiv, (size_t)2 * L * NT) ||
            dalloc(&c->d_icy, (size_t)L * Ni * 7) || dalloc(&c->d_icblk, icblocks.size()) ||
            dalloc(&c->d_icord, H) || dalloc(&c->d_icpartial, (size_t)4 * c->n_icblk))
            return LSD_E_HIP;
        for (int k = 0; k < 2; k++)
            if (dalloc(&c->d_icf[k], (size_t)H * Ni) || dalloc(&c->d_icn[k], (size_t)H * Ni) ||
                dalloc(&c->d_ichist[k], (size_t)H * Ni))
                return LSD_E_HIP;
        LSD_HIP(hipHostMalloc((void**)&c->h_icpartial, sizeof(double) * 4 * c->n_icblk, hipHostMallocDefault));
    }

    hipStream_t st = c->stream;
    LSD_HIP(hipMemcpyAsync(c->d_hl, c->hl.data(), sizeof(int) * H, hipMemcpyHostToDevice, st));
    LSD_HIP(hipMemcpyAsync(c->d_hm, c->hm.data(), sizeof(int) * H, hipMemcpyHostToDevice, st));
    LSD_HIP(hipMemcpyAsync(c->d_pidx, pidx.data(), sizeof(int) * pidx.size(), hipMemcpyHostToDevice, st));
    LSD_HIP(hipMemcpyAsync(c->d_apos, apos.data(), sizeof(int) * H, hipMemcpyHostToDevice, st));
    LSD_HIP(hipMemsetAsync(c->d_wq, 0, sizeof(int) * 32, st));
    if (c->Nic > 1) {
        LSD_HIP(hipMemcpyAsync(c->d_icord, icorder.data(), sizeof(int) * H, hipMemcpyHostToDevice, st));
        LSD_HIP(hipMemcpyAsync(c->d_icblk, icblocks.data(), sizeof(int4) * icblocks.size(), hipMemcpyHostToDevice,
                               st));
        for (int k = 0; k < 2; k++) {
            LSD_HIP(hipMemsetAsync(c->d_icf[k], 0, sizeof(double2) * H * c->Nic, st));
            LSD_HIP(hipMemsetAsync(c->d_icn[k], 0, sizeof(double2) * H * c->Nic, st));
            LSD_HIP(hipMemsetAsync(c->d_ichist[k], 0, sizeof(double2) * H * c->Nic, st));
        }
        LSD_HIP(hipMemsetAsync(c->d_iclu, 0, sizeof(double) * 2 * L * c->NT * LSD_LDAB, st));
        LSD_HIP(hipMemsetAsync(c->d_icpiv, 0, sizeof(int) * 2 * L * c->NT, st));
        LSD_HIP(hipMemsetAsync(c->d_icy, 0, sizeof(double) * L * c->Nic * 7, st));
    }
    for (int form = 0; form < 3; form++) {
        if (!c->d_sblk[form]) continue;
        LSD_HIP(hipMemcpyAsync(c->d_sord[form], sorder[form].data(), sizeof(int) * H, hipMemcpyHostToDevice, st));
        LSD_HIP(hipMemcpyAsync(c->d_sblk[form], sblocks[form].data(), sizeof(int4) * sblocks[form].size(),
                               hipMemcpyHostToDevice, st));
    }
    LSD_HIP(hipMemcpyAsync(c->d_slots, c->slots.data(), sizeof(Slot) * c->nslot, hipMemcpyHostToDevice, st));
    LSD_HIP(hipMemcpyAsync(c->d_slot_order, c->slot_order.data(), sizeof(int) * c->nslot, hipMemcpyHostToDevice, st));
    {
        std::vector<int> mslot(c->M, 0);
        for (int k = 0; k < c->nslot; k++) mslot[c->slots[k].mfft] = k;
        LSD_HIP(hipMemcpyAsync(c->d_mslot, mslot.data(), sizeof(int) * c->M, hipMemcpyHostToDevice, st));
        LSD_HIP(hipStreamSynchronize(st));
    }
    LSD_HIP(hipMemsetAsync(c->d_tab, 0, sizeof(double) * 5 * H * c->Th2, st));
    LSD_HIP(hipMemsetAsync(c->d_atab, 0, sizeof(double) * 3 * H * c->Th2, st));
    LSD_HIP(hipMemsetAsync(c->d_lu, 0, sizeof(double) * NLU * L * N * LSD_LDAB, st));
    LSD_HIP(hipMemsetAsync(c->d_piv, 0, sizeof(int) * NLU * L * N, st));
    LSD_HIP(hipMemsetAsync(c->d_y, 0, sizeof(double) * NY * L * N * 7, st));
    LSD_HIP(hipMemsetAsync(c->d_xband, 0, sizeof(double) * NLU * L * N * 7, st));
    LSD_HIP(hipMemsetAsync(c->d_gre, 0, sizeof(double) * 2 * L * N, st));
    LSD_HIP(hipMemsetAsync(c->d_invg, 0, sizeof(double) * 4 * L, st));
    for (int k = 0; k < LSD_NUM_COLL; k++) LSD_HIP(hipMemsetAsync(c->d_f[k], 0, sizeof(double2) * HN, st));
    for (int k = 0; k < 6; k++) LSD_HIP(hipMemsetAsync(c->d_hist[k], 0, sizeof(double2) * HN, st));
    LSD_HIP(hipMemsetAsync(c->d_spec_in, 0, sizeof(double2) * SI_COUNT * HN, st));
    LSD_HIP(hipMemsetAsync(c->d_spec_out, 0, sizeof(double2) * SO_COUNT * HN, st));
    LSD_HIP(hipMemsetAsync(c->d_fourier, 0, sizeof(double2) * LSD_NUM_PHYS * c->fourier_elems(), st));
    LSD_HIP(hipMemsetAsync(c->d_folded, 0, sizeof(double2) * NO_COUNT * c->folded_elems(), st));
    LSD_HIP(hipMemsetAsync(c->d_phys, 0, sizeof(double) * LSD_NUM_PHYS * c->phys_elems(), st));
    LSD_HIP(hipMemsetAsync(c->d_nlphys, 0, sizeof(double) * NO_COUNT * c->phys_elems(), st));

    if ((rc = fft_setup(*c))) return rc;
    if ((rc = precompute_mesh(*c))) return rc;
    if (dalloc(&c->d_cfl_d, (size_t)2 * N) || dalloc(&c->d_cfl_part, (size_t)2 * N * std::max(c->Th2, 1)) ||
        dalloc(&c->d_cfl, 1))
        return LSD_E_HIP;
    LSD_HIP(hipHostMalloc((void**)&c->h_cfl, sizeof(double), hipHostMallocDefault));
    if ((rc = cfl_setup(*c))) return rc;
    if ((rc = precompute_tables(*c))) return rc;
    if (c->mixed) {
        const size_t HT = (size_t)H * c->Th2;
        c->Th2p = (c->Th2 + 63) / 64 * 64;
        std::vector<int> t0(c->nslot), row_h;
        for (int s = 0; s < c->nslot; s++) {
            t0[s] = (int)row_h.size();
            const int len = c->slots[s].len, rows = (len + 7) / 8 * 8;
            for (int k = 0; k < rows; k++) row_h.push_back(k < len ? c->slots[s].h0 + k : -1);
        }
        c->Hp = (int)row_h.size();
        if (dalloc(&c->d_tab32, (size_t)5 * c->Hp * c->Th2p) || dalloc(&c->d_atab32, 3 * HT) ||
            dalloc(&c->d_slot_t0, c->nslot) || dalloc(&c->d_row_h, c->Hp))
            return LSD_E_HIP;
        LSD_HIP(hipMemcpyAsync(c->d_slot_t0, t0.data(), sizeof(int) * c->nslot, hipMemcpyHostToDevice, st));
        LSD_HIP(hipMemcpyAsync(c->d_row_h, row_h.data(), sizeof(int) * c->Hp, hipMemcpyHostToDevice, st));
        LSD_HIP(hipStreamSynchronize(st));
        if ((rc = tables_to_f32(*c))) return rc;
        // wide-load analysis tables (even Th2 only: its loads are 16 bytes)
        if (c->Th2 % 2 == 0) {
            std::vector<int> base(2 * c->nslot), acol;
            int maxhb = 0;
            for (int sl = 0; sl < c->nslot; sl++) {
                const Slot& q = c->slots[sl];
                const int n_even = (q.len + 1) / 2;
                for (int par = 0; par < 2; par++) {
                    const int ncls = par ? q.len / 2 : n_even;
                    const int nblk = (ncls + 63) / 64;
                    maxhb = std::max(maxhb, nblk);
                    base[2 * sl + par] = (int)acol.size();
                    for (int w = 0; w < 64 * nblk; w++) {
                        const int hb = w / 64, r = w % 64, kc = 64 * hb + 16 * (r & 3) + (r >> 2);
                        acol.push_back(kc < ncls ? q.h0 + (par ? n_even : 0) + kc : -1);
                    }
                }
            }
            c->Hpa = (int)acol.size();
            c->ana_maxhb = maxhb;
            c->Th2a = (c->Th2 + 7) / 8 * 8;
            if (dalloc(&c->d_atab4, (size_t)3 * c->Th2a * c->Hpa) || dalloc(&c->d_ana_base, 2 * c->nslot) ||
                dalloc(&c->d_acol, c->Hpa))
                return LSD_E_HIP;
            LSD_HIP(hipMemcpyAsync(c->d_ana_base, base.data(), sizeof(int) * 2 * c->nslot, hipMemcpyHostToDevice, st));
            LSD_HIP(hipMemcpyAsync(c->d_acol, acol.data(), sizeof(int) * c->Hpa, hipMemcpyHostToDevice, st));
            LSD_HIP(hipStreamSynchronize(st));
            if ((rc = tables_ana4(*c))) return rc;
        }
    }
    if (cfg->precision == LSD_MIXED_BF16X3 && (rc = tables_bf16(*c))) return rc;
    c->ana_wide_ok = c->d_atab4 != nullptr;
    if ((rc = precompute_innercore(*c))) return rc;
    if ((rc = build_operators(*c))) return rc;
    LSD_HIP(hipStreamSynchronize(st));
    load_tuning(*c);
    return LSD_OK;
}

// ------------------------------------------------------------ timing

ProfScope::ProfScope(Ctx& ctx, int cls, double work) : c(ctx), on(ctx.prof_class == cls) {
    if (!on) return;
    hipEvent_t start;
    if (c.cap >= 0) {
        // inside a capture: a pair of its own, replayed with the graph
        if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
            on = false;
            return;
        }
        c.g_ev[c.cap].push_back(start);
        c.g_ev[c.cap].push_back(stop);
        c.g_work[c.cap] += work;
    } else {
        if ((size_t)(c.prof_used + 2) > c.prof_events.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) {
                    on = false;
                    return;
                }
                c.prof_events.push_back(e);
            }
        }
        start = c.prof_events[c.prof_used];
        stop = c.prof_events[c.prof_used + 1];
        c.prof_used += 2;
        c.prof_work += work;
    }
    (void)hipEventRecord(start, c.stream);
}

ProfScope::~ProfScope() {
    if (on && stop) (void)hipEventRecord(stop, c.stream);
}

// After a replay of graph `which` has completed: add its timed launches.
static int graph_prof_accumulate(Ctx& c, int which) {
    if (c.prof_class < 0 || c.g_prof_class != c.prof_class) return LSD_OK;
    const std::vector<hipEvent_t>& ev = c.g_ev[which];
    for (size_t k = 0; k + 1 < ev.size(); k += 2) {
        float ms = 0.f;
        LSD_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        c.prof_acc_ms += ms;
        c.prof_acc_launches++;
    }
    c.prof_acc_work += c.g_work[which];
    return LSD_OK;
}

static void drop_graphs(Ctx& c) {
    if (c.g_first) (void)hipGraphExecDestroy(c.g_first);
    if (c.g_more) (void)hipGraphExecDestroy(c.g_more);
    if (c.g_eval) (void)hipGraphExecDestroy(c.g_eval);
    c.g_first = c.g_more = c.g_eval = nullptr;
    for (int w = 0; w < 2; w++) {
        for (hipEvent_t e : c.g_ev[w]) (void)hipEventDestroy(e);
        c.g_ev[w].clear();
        c.g_work[w] = 0.0;
    }
    c.g_prof_class = -1;
}

// ------------------------------------------------------------- phases

static double2* spec_in(Ctx& c, int k) { return c.d_spec_in + (size_t)k * c.coll_elems(); }
static double2* spec_out(Ctx& c, int k) { return c.d_spec_out + (size_t)k * c.coll_elems(); }
static double2* fourier(Ctx& c, int k) { return c.d_fourier + (size_t)k * c.fourier_elems(); }
static double2* folded(Ctx& c, int k) { return c.d_folded + (size_t)k * c.folded_elems(); }
static double* phys(Ctx& c, int k) { return c.d_phys + (size_t)k * c.phys_elems(); }
static double* nlphys(Ctx& c, int k) { return c.d_nlphys + (size_t)k * c.phys_elems(); }

#define RC(x)             \
    do {                  \
        int r_ = (x);     \
        if (r_) return r_; \
    } while (0)

// Buffers in the transform precision: FP32 forms use the first half of the
// FP64 allocations.
template <typename B>
static void* slot_p(B* base, size_t per, int k, bool f32) {
    if (f32) return reinterpret_cast<float2*>(base) + (size_t)k * per;
    return reinterpret_cast<double2*>(base) + (size_t)k * per;
}
static void* spec_in_p(Ctx& c, int k, bool f) { return slot_p(c.d_spec_in, c.coll_elems(), k, f); }
static void* spec_out_p(Ctx& c, int k, bool f) { return slot_p(c.d_spec_out, c.coll_elems(), k, f); }
static void* fourier_p(Ctx& c, int k, bool f) { return slot_p(c.d_fourier, c.fourier_elems(), k, f); }
static void* folded_p(Ctx& c, int k, bool f) { return slot_p(c.d_folded, c.folded_elems(), k, f); }

// Synthesis of the fields the nonlinear terms need (vel_transform,
// mag_transform, cod_transform, comp_transform): the spectral preparations
// first, then one launch per transform kind over all the fields of that
// kind, then (when `to_phys`, FP64 only) the c2r FFTs to the physical
// buffers.
static int synthesise(Ctx& c, bool f, bool to_phys) {
    const void* sc_in[LG_SCALARS];
    void* sc_out[LG_SCALARS];
    const void* v_s[4];
    const void* v_t[4];
    void* v_th[4];
    void* v_ph[4];
    const void* g_in[2];
    void* g_th[2];
    void* g_ph[2];
    int ns = 0, nv = 0, ng = 0;
    // B and curl B of the LSD_MIXED_F32_MAG64 form: FP64 preparation and
    // Legendre sums, FP32 Fourier outputs (two scalar, two vector jobs)
    const bool m64 = f && c.mag64;
    const void* m_sc_in[2];
    void* m_sc_out[2];
    const void* m_v_s[2];
    const void* m_v_t[2];
    void* m_v_th[2];
    void* m_v_ph[2];
    // vel_angmomclr before the velocity transform (program/main.F90:100)
    RC(launch_angmomclr(c));
    // the vector fields (u; B) and the scalar fields (C; xi) are prepared in
    // one launch each
    const double2* vT[2];
    const double2* vP[2];
    void* vo[2][6];
    int nvf = 0;
    auto vec = [&](int T, int P, int si, int pr) {
        void** o = vo[nvf];
        // the FP32 slots fill the first half of the allocation, so the FP64
        // inputs of B and curl B take FP64 slots from the second half
        const bool own = m64 && T == LSD_MAG_TOR;
        static_assert(SI_COUNT / 2 + 6 <= SI_COUNT, "six FP64 slots in the second half");
        for (int k = 0; k < 6; k++) o[k] = own ? spec_in_p(c, SI_COUNT / 2 + k, false) : spec_in_p(c, si + k, f);
        vT[nvf] = c.d_f[T];
        vP[nvf++] = c.d_f[P];
        if (own) {
            for (int u = 0; u < 2; u++) {
                m_sc_in[u] = o[3 * u];
                m_sc_out[u] = fourier_p(c, pr + 3 * u, true);
                m_v_s[u] = o[3 * u + 1];
                m_v_t[u] = o[3 * u + 2];
                m_v_th[u] = fourier_p(c, pr + 3 * u + 1, true);
                m_v_ph[u] = fourier_p(c, pr + 3 * u + 2, true);
            }
            return;
        }
        // q of the field and of its curl (scalar), s and t of both (vector)
        sc_in[ns] = o[0];
        sc_out[ns++] = fourier_p(c, pr, f);
        sc_in[ns] = o[3];
        sc_out[ns++] = fourier_p(c, pr + 3, f);
        v_s[nv] = o[1];
        v_t[nv] = o[2];
        v_th[nv] = fourier_p(c, pr + 1, f);
        v_ph[nv++] = fourier_p(c, pr + 2, f);
        v_s[nv] = o[4];
        v_t[nv] = o[5];
        v_th[nv] = fourier_p(c, pr + 4, f);
        v_ph[nv++] = fourier_p(c, pr + 5, f);
    };
    vec(LSD_VEL_TOR, LSD_VEL_POL, SI_UQ, LSD_P_VEL_R);
    if (c.cfg.mag) vec(LSD_MAG_TOR, LSD_MAG_POL, SI_BQ, LSD_P_MAG_R);
    if (m64 && nvf == 2) {
        void* const* po[2] = {vo[0], vo[1]};
        RC(launch_vec_prep(c, 1, vT, vP, po, true));
        RC(launch_vec_prep(c, 1, vT + 1, vP + 1, po + 1, false));
    } else {
        void* const* po[2] = {vo[0], vo[1]};
        RC(launch_vec_prep(c, nvf, vT, vP, po, f));
    }
    // scalars: radial derivative (scalar synthesis), the field itself
    // (angular gradient)
    const int sf[2] = {LSD_COD, LSD_COMP}, sd[2] = {SI_DC, SI_DX}, scc[2] = {SI_CC, SI_XC};
    const int pr[2] = {LSD_P_COD_GR, LSD_P_COMP_GR};
    const double2* sC[2];
    void* sdC[2];
    void* sCc[2];
    int nsf = 0;
    for (int q = 0; q < 2; q++) {
        if (!(q == 0 ? c.cfg.cod : c.cfg.comp)) continue;
        void* dC = spec_in_p(c, sd[q], f);
        void* Cc = f ? spec_in_p(c, scc[q], f) : nullptr;
        sC[nsf] = c.d_f[sf[q]];
        sdC[nsf] = dC;
        sCc[nsf++] = Cc;
        sc_in[ns] = dC;
        sc_out[ns++] = fourier_p(c, pr[q], f);
        g_in[ng] = f ? Cc : (void*)c.d_f[sf[q]];
        g_th[ng] = fourier_p(c, pr[q] + 1, f);
        g_ph[ng++] = fourier_p(c, pr[q] + 2, f);
    }
    if (nsf) RC(launch_scal_prep(c, nsf, sC, sdC, sCc, f));
    RC(launch_synthesis_group(c, SYN_SCALAR, ns, sc_in, nullptr, sc_out, nullptr, f));
    RC(launch_synthesis_group(c, SYN_VECTOR, nv, v_s, v_t, v_th, v_ph, f));
    if (ng) RC(launch_synthesis_group(c, SYN_GRAD, ng, g_in, nullptr, g_th, g_ph, f));
    if (m64 && nvf == 2) {
        RC(launch_synthesis_group(c, SYN_SCALAR, 2, m_sc_in, nullptr, m_sc_out, nullptr, false, true));
        RC(launch_synthesis_group(c, SYN_VECTOR, 2, m_v_s, m_v_t, m_v_th, m_v_ph, false, true));
    }
    // non_rotation after the transforms (program/main.F90:108)
    RC(launch_torques(c, f));
    if (to_phys && !f) {
        RC(launch_fft_c2r(c, fourier(c, 0), phys(c, 0), 6));
        if (c.cfg.mag) RC(launch_fft_c2r(c, fourier(c, 6), phys(c, 6), 6));
        for (int q = 0; q < 2; q++)
            if (q == 0 ? c.cfg.cod : c.cfg.comp) RC(launch_fft_c2r(c, fourier(c, pr[q]), phys(c, pr[q]), 3));
    }
    return LSD_OK;
}

// vel_transform, cod_transform, comp_transform, mag_transform (FP64, to the
// physical buffers)
static int transform(Ctx& c) { return synthesise(c, false, true); }

// Analysis of the folded nonlinear products and the coll-side completion of
// non_velocity, non_codensity, non_composition, non_magnetic.
static int analyse(Ctx& c, bool f) {
    const void* v0[2];
    const void* v1[2];
    const void* v2[2];
    void* q[2];
    void* s_[2];
    void* t[2];
    const void* sg[2];
    void* so[2];
    int nv = 0, ns = 0;
    v0[nv] = folded_p(c, NO_UR, f);
    v1[nv] = folded_p(c, NO_UT, f);
    v2[nv] = folded_p(c, NO_UP, f);
    q[nv] = spec_out_p(c, SO_UQ, f);
    s_[nv] = spec_out_p(c, SO_US, f);
    t[nv++] = spec_out_p(c, SO_UT, f);
    if (c.cfg.mag) {
        v0[nv] = folded_p(c, NO_ER, f);
        v1[nv] = folded_p(c, NO_ET, f);
        v2[nv] = folded_p(c, NO_EP, f);
        q[nv] = spec_out_p(c, SO_EQ, f);
        s_[nv] = spec_out_p(c, SO_ES, f);
        t[nv++] = spec_out_p(c, SO_ET, f);
    }
    if (c.cfg.cod) {
        sg[ns] = folded_p(c, NO_C, f);
        so[ns++] = spec_out_p(c, SO_C, f);
    }
    if (c.cfg.comp) {
        sg[ns] = folded_p(c, NO_X, f);
        so[ns++] = spec_out_p(c, SO_X, f);
    }
    RC(launch_analysis_group(c, ANA_VECTOR, nv, v0, v1, v2, q, s_, t, f));
    if (ns) RC(launch_analysis_group(c, ANA_SCALAR, ns, sg, nullptr, nullptr, so, nullptr, nullptr, f));
    RC(launch_spec_post(c, f));
    RC(launch_ic_nonlinear(c));
    return LSD_OK;
}

// non_velocity, non_codensity, non_composition, non_magnetic on the
// physical buffers left by transform() (FP64)
static int nonlinear(Ctx& c) {
    RC(launch_nonlinear(c));
    RC(launch_fft_r2c_fold(c, nlphys(c, NO_UR), folded(c, NO_UR), 3));
    if (c.cfg.cod) RC(launch_fft_r2c_fold(c, nlphys(c, NO_C), folded(c, NO_C), 1));
    if (c.cfg.comp) RC(launch_fft_r2c_fold(c, nlphys(c, NO_X), folded(c, NO_X), 1));
    if (c.cfg.mag) RC(launch_fft_r2c_fold(c, nlphys(c, NO_ER), folded(c, NO_ER), 3));
    return analyse(c, false);
}

// One nonlinear evaluation inside the timestep: synthesis, the fused
// physical-space pass, analysis.  The physical fields stay on chip.
static int evaluate(Ctx& c) {
    if (!physfuse_supported(c)) {
        RC(transform(c));
        if (c.cfl_on) RC(launch_cfl_phys(c));
        return nonlinear(c);
    }
    RC(synthesise(c, c.mixed, false));
    RC(launch_physfuse(c, c.mixed));
    return analyse(c, c.mixed);
}

// tim_check_cgce (modules/timestep.F90:744-830).  Control 2 iterates while
// tim_dterr > d_dterr and estimates tim_corr_dt = dt sqrt(d_dterr /
// tim_dterr); control 1 iterates while tim_dterr > 1 and takes tim_corr_dt =
// tim_dterr.  Returns nonzero if the reference would stop the run.
static int check_cgce(Ctx& c) {
    int stop = 0;
    const bool ctl1 = c.cfg.step_control == 1;
    const double crit = ctl1 ? 1.0 : c.cfg.dterr;
    if (c.it == 1) {
        c.corr_dt = ctl1 ? c.dterr : c.dt * std::sqrt(c.cfg.dterr / c.dterr);
        c.lasterr = 1e99;
    }
    if (c.dt < 1e-12 && c.step > 30) {
        stop = 1;
        c.it = 0;
    } else if (c.it > 10) {
        stop = 1;
        c.it = 0;
    } else if (c.dterr > c.lasterr) {
        if (ctl1) {
            if (c.dterr > 1.0) stop = 1;
        } else {
            if (c.dterr > 2.0 * c.cfg.dterr) stop = 1;
            if (c.dterr < 2.0 * c.cfg.dterr) c.corr_dt = c.dt / (1e1 * COURANT);
        }
        c.it = 0;
    } else if (c.dterr > crit) {
        c.lasterr = c.dterr;
        c.it = c.it + 1;
    } else {
        c.it = 0;
    }
    c.dterr = 0.0;
    return stop;
}

// Host side of one corrector pass: accumulate tim_dterr, tim_check_cgce.
static int after_corrector(Ctx& c, int& its, double& last) {
    c.dterr += corrector_norms(c);
    last = c.dterr;
    its++;
    if (check_cgce(c)) {
        set_error("tim_check_cgce: run stopped");
        return LSD_E_DIVERGED;
    }
    return LSD_OK;
}

// Graph kinds: CAP_FIRST [evaluate, predictor, evaluate, corrector], CAP_MORE
// [evaluate, corrector] (fixed dt); with the adaptive step the first pass
// splits at the step choice: CAP_EVAL [evaluate with the CFL partials,
// tim_cfl_dt] and CAP_REST [predictor, evaluate, corrector].  Each corrector
// ends with the copy of its norm partials.
enum CaptureKind { CAP_FIRST = 0, CAP_MORE = 1, CAP_EVAL = 2, CAP_REST = 3 };

static int capture(Ctx& c, int kind, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    LSD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.cap = kind == CAP_MORE ? 1 : 0;
    int rc = LSD_OK;
    if (kind == CAP_EVAL) {
        c.cfl_on = true;
        rc = evaluate(c);
        c.cfl_on = false;
        if (!rc) rc = launch_cfl_final(c);
    } else {
        if (kind != CAP_REST) rc = evaluate(c);
        if (!rc && kind != CAP_MORE) rc = launch_predictor(c);
        if (!rc && kind != CAP_MORE) rc = evaluate(c);
        if (!rc) rc = launch_corrector_async(c);
    }
    c.cap = -1;
    hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    LSD_HIP(e);
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    LSD_HIP(ie);
    return LSD_OK;
}

static int ensure_graphs(Ctx& c) {
    if (c.g_first && c.g_prof_class != c.prof_class) drop_graphs(c);
    if (!c.g_first) {
        RC(capture(c, c.adaptive ? CAP_REST : CAP_FIRST, &c.g_first));
        RC(capture(c, CAP_MORE, &c.g_more));
        if (c.adaptive) RC(capture(c, CAP_EVAL, &c.g_eval));
        c.g_prof_class = c.prof_class;
    }
    return LSD_OK;
}

// Adaptive step, after the predictor pass's evaluation has queued
// tim_cfl_dt: non_maxtstep's result, tim_new_tstep, and the *_matrices
// rebuild on a new dt (program/main.F90:124-135).
static int choose_dt(Ctx& c) {
    LSD_HIP(hipStreamSynchronize(c.stream));
    c.cfl_dt = *c.h_cfl;
    if (new_tstep(c)) RC(build_operators(c));
    return LSD_OK;
}

// main.F90:100-164 replayed from the captured graphs: g_first covers the
// predictor pass and the first corrector pass of a step (with the adaptive
// step, g_eval first takes the predictor pass's evaluation up to the step
// choice), g_more any further corrector pass tim_check_cgce asks for.
static int step_graph(Ctx& c, int nsteps, lsd_step_stats* z) {
    RC(ensure_graphs(c));
    for (int k = 0; k < nsteps; k++) {
        int its = 0;
        double last = 0.0;
        if (c.adaptive) {
            LSD_HIP(hipGraphLaunch(c.g_eval, c.stream));
            RC(choose_dt(c));
        }
        LSD_HIP(hipGraphLaunch(c.g_first, c.stream));
        LSD_HIP(hipStreamSynchronize(c.stream));
        RC(graph_prof_accumulate(c, 0));
        z->evaluations += 2;
        c.it = 1;
        RC(after_corrector(c, its, last));
        while (c.it != 0) {
            LSD_HIP(hipGraphLaunch(c.g_more, c.stream));
            LSD_HIP(hipStreamSynchronize(c.stream));
            RC(graph_prof_accumulate(c, 1));
            z->evaluations++;
            RC(after_corrector(c, its, last));
        }
        c.t = c.t + c.dt;
        c.step++;
        z->steps++;
        z->last_its = its;
        z->max_its = std::max(z->max_its, its);
        z->last_dterr = last;
    }
    return LSD_OK;
}

static int step(Ctx& c, int nsteps, lsd_step_stats* st) {
    lsd_step_stats z;
    std::memset(&z, 0, sizeof z);
    // graphs need the loop to start a fresh step; HIP events recorded inside
    // a graph cannot be timed, so per-class timing runs the eager loop
    if (c.use_graph && c.it == 0 && c.prof_class < 0) {
        RC(step_graph(c, nsteps, &z));
    } else {
        for (int k = 0; k < nsteps; k++) {
            int its = 0;
            double last = 0.0;
            do {
                c.cfl_on = c.adaptive && c.it == 0;
                int rc = evaluate(c);
                c.cfl_on = false;
                if (rc) return rc;
                z.evaluations++;
                if (c.it == 0) {
                    if (c.adaptive) {
                        RC(launch_cfl_final(c));
                        RC(choose_dt(c));
                    }
                    RC(launch_predictor(c));
                    c.it = 1;
                } else {
                    RC(launch_corrector_async(c));
                    LSD_HIP(hipStreamSynchronize(c.stream));
                    RC(after_corrector(c, its, last));
                }
            } while (c.it != 0);
            c.t = c.t + c.dt;
            c.step++;
            z.steps++;
            z.last_its = its;
            z.max_its = std::max(z.max_its, its);
            z.last_dterr = last;
        }
    }
    z.t = c.t;
    if (st) *st = z;
    return LSD_OK;
}

// Steps independent contexts together: each pass launches every member's
// graph on its own stream before waiting for any, so the members run
// concurrently; each member's tim_check_cgce then runs on the host.  A
// member that fails is drained and dropped from the remaining passes; the
// others finish all nsteps.
static int step_many(std::vector<Ctx*>& cs, int nsteps, lsd_step_stats* st) {
    const size_t n = cs.size();
    std::vector<lsd_step_stats> z(n);
    for (auto& q : z) std::memset(&q, 0, sizeof q);
    for (Ctx* c : cs) {
        if (!c->use_graph || c->it != 0 || c->prof_class >= 0) {
            set_error("lsd_step_many: every member must be between steps, with graphs on and no profiling");
            return LSD_E_ARG;
        }
        if (c->adaptive) {
            set_error("lsd_step_many: members must have a fixed dt");
            return LSD_E_ARG;
        }
        LSD_HIP(hipSetDevice(c->device));
        RC(ensure_graphs(*c));
    }
    std::vector<int> its(n), more(n), status(n, LSD_OK);
    std::vector<double> last(n);
    std::vector<std::string> why(n);
    // member i stops with code rc: its queued work is waited for (errors of
    // the wait are already reported through rc)
    auto fail = [&](size_t i, int rc) {
        status[i] = rc;
        why[i] = g_err;
        more[i] = 0;
        (void)hipSetDevice(cs[i]->device);
        (void)hipStreamSynchronize(cs[i]->stream);
    };
    for (int k = 0; k < nsteps; k++) {
        for (size_t i = 0; i < n; i++) {
            its[i] = 0;
            more[i] = status[i] == LSD_OK;
            if (!more[i]) continue;
            hipError_t e = hipSetDevice(cs[i]->device);
            if (e == hipSuccess) e = hipGraphLaunch(cs[i]->g_first, cs[i]->stream);
            if (e != hipSuccess) {
                set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
                fail(i, LSD_E_HIP);
            }
        }
        bool pending = true;
        while (pending) {
            pending = false;
            for (size_t i = 0; i < n; i++) {
                if (!more[i]) continue;
                Ctx& c = *cs[i];
                hipError_t e = hipSetDevice(c.device);
                if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
                if (e != hipSuccess) {
                    set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
                    fail(i, LSD_E_HIP);
                    continue;
                }
                if (its[i] == 0) {
                    z[i].evaluations += 2;
                    c.it = 1;
                } else {
                    z[i].evaluations++;
                }
                int rc = after_corrector(c, its[i], last[i]);
                if (rc) {
                    fail(i, rc);
                    continue;
                }
                if (c.it != 0) {
                    e = hipGraphLaunch(c.g_more, c.stream);
                    if (e != hipSuccess) {
                        set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
                        fail(i, LSD_E_HIP);
                        continue;
                    }
                    pending = true;
                } else {
                    more[i] = 0;
                    c.t = c.t + c.dt;
                    c.step++;
                    z[i].steps++;
                    z[i].last_its = its[i];
                    z[i].max_its = std::max(z[i].max_its, its[i]);
                    z[i].last_dterr = last[i];
                }
            }
        }
    }
    int first = LSD_OK;
    std::string msg;
    for (size_t i = 0; i < n; i++) {
        z[i].t = cs[i]->t;
        z[i].status = status[i];
        if (st) st[i] = z[i];
        if (status[i] != LSD_OK) {
            if (first == LSD_OK) first = status[i];
            msg += (msg.empty() ? "" : "; ") + std::string("member ") + std::to_string(i) + " after " +
                   std::to_string(z[i].steps) + " steps: " + why[i];
        }
    }
    if (first != LSD_OK) set_error("lsd_step_many: " + msg);
    return first;
}

// ------------------------------------------------------------ autotuner

// Average device time in microseconds of `reps` runs of body on the
// context's stream (after one untimed run).
template <class F>
static int time_us(Ctx& c, int reps, F&& body, double* out) {
    RC(body());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    LSD_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        LSD_HIP(he);
    }
    int rc = LSD_OK;
    if (hipEventRecord(e0, c.stream) != hipSuccess) rc = LSD_E_HIP;
    for (int k = 0; k < reps && rc == LSD_OK; k++) rc = body();
    float ms = 0.f;
    if (rc == LSD_OK && (hipEventRecord(e1, c.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
        set_error("lsd_tune: event timing failed");
        rc = LSD_E_HIP;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *out = 1e3 * (double)ms / reps;
    return LSD_OK;
}

static int tune(Ctx& c, int reps, const char* cache, double* tus) {
    if (c.it != 0 || c.prof_class >= 0) {
        set_error("lsd_tune: only between timesteps and without profiling");
        return LSD_E_ARG;
    }
    reps = std::max(reps, 1);
    double t[LSD_V_COUNT][2] = {};
    drop_graphs(c);
    // FP32 synthesis shapes and analysis forms: the launches of one
    // evaluation, on the current buffers (they only write derived buffers)
    if (c.mixed) {
        for (int k = LSD_V_SYN_SCALAR; k <= LSD_V_SYN_GRAD; k++) {
            for (int v = 0; v < 2; v++) {
                c.var[k] = v ? 8 : 4;
                RC(time_us(c, reps, [&]() { return synthesise(c, true, false); }, &t[k][v]));
            }
            c.var[k] = t[k][0] <= t[k][1] ? 4 : 8;
            // the two persistent forms of the faster shape (plain, and with
            // the next item's loads issued before each epilogue)
            const int best = c.var[k];
            int choice = best;
            double tbest = std::min(t[k][0], t[k][1]);
            for (int d = 1; d <= 2; d++) {
                double tp = 0.0;
                c.var[k] = best + d;
                RC(time_us(c, reps, [&]() { return synthesise(c, true, false); }, &tp));
                if (tp < tbest) {
                    tbest = tp;
                    choice = best + d;
                }
            }
            c.var[k] = choice;
        }
        if (c.ana_wide_ok) {
            for (int v = 0; v < 2; v++) {
                c.var[LSD_V_ANA] = v;
                RC(time_us(c, reps, [&]() { return analyse(c, true); }, &t[LSD_V_ANA][v]));
            }
            c.var[LSD_V_ANA] = t[LSD_V_ANA][0] <= t[LSD_V_ANA][1] ? 0 : 1;
            if (c.var[LSD_V_ANA] == 0) {
                // the persistent form of the wide-load kernel
                double tp = 0.0;
                c.var[LSD_V_ANA] = 2;
                RC(time_us(c, reps, [&]() { return analyse(c, true); }, &tp));
                if (tp >= t[LSD_V_ANA][0]) c.var[LSD_V_ANA] = 0;
            }
        }
    }
    // solve forms: predictor + corrector on the current state, which is
    // saved first and restored afterwards
    if (c.solve_direct_ok && c.hb_form[1] >= 1) {
        const size_t HN = c.coll_elems();
        double2* save = nullptr;
        AsyncScratch sv(c.stream);
        LSD_HIP(sv.alloc(&save, sizeof(double2) * 12 * HN));
        for (int k = 0; k < 6; k++) {
            LSD_HIP(hipMemcpyAsync(save + k * HN, c.d_f[k], sizeof(double2) * HN, hipMemcpyDeviceToDevice, c.stream));
            LSD_HIP(hipMemcpyAsync(save + (6 + k) * HN, c.d_hist[k], sizeof(double2) * HN, hipMemcpyDeviceToDevice,
                                   c.stream));
        }
        for (int v = 0; v < 2; v++) {
            c.var[LSD_V_SOLVE] = v;
            select_solve_form(c);
            RC(time_us(
                c, reps,
                [&]() {
                    RC(launch_predictor(c));
                    return launch_corrector_async(c);
                },
                &t[LSD_V_SOLVE][v]));
        }
        c.var[LSD_V_SOLVE] = t[LSD_V_SOLVE][0] <= t[LSD_V_SOLVE][1] ? 0 : 1;
        select_solve_form(c);
        for (int k = 0; k < 6; k++) {
            LSD_HIP(hipMemcpyAsync(c.d_f[k], save + k * HN, sizeof(double2) * HN, hipMemcpyDeviceToDevice, c.stream));
            LSD_HIP(hipMemcpyAsync(c.d_hist[k], save + (6 + k) * HN, sizeof(double2) * HN, hipMemcpyDeviceToDevice,
                                   c.stream));
        }
        LSD_HIP(hipStreamSynchronize(c.stream));
    }
    if (tus)
        for (int k = 0; k < LSD_V_COUNT; k++) {
            tus[2 * k] = t[k][0];
            tus[2 * k + 1] = t[k][1];
        }
    if (cache) {
        FILE* f = std::fopen(cache, "a");
        if (!f) {
            set_error(std::string("lsd_tune: cannot open ") + cache);
            return LSD_E_ARG;
        }
        std::fprintf(f,
                     "{\"L\": %d, \"M\": %d, \"Mp\": %d, \"N\": %d, \"Th\": %d, \"Ph\": %d, \"precision\": %d, "
                     "\"syn_scalar\": %d, \"syn_vector\": %d, \"syn_grad\": %d, \"ana\": %d, \"solve\": %d, "
                     "\"us\": [%.1f, %.1f, %.1f, %.1f, %.1f, %.1f, %.1f, %.1f, %.1f, %.1f]}\n",
                     c.L, c.M, c.Mp, c.N, c.Th, c.Ph, c.cfg.precision, c.var[0], c.var[1], c.var[2], c.var[3],
                     c.var[4], t[0][0], t[0][1], t[1][0], t[1][1], t[2][0], t[2][1], t[3][0], t[3][1], t[4][0],
                     t[4][1]);
        if (std::fclose(f) != 0) {
            set_error(std::string("lsd_tune: write failed for ") + cache);
            return LSD_E_ARG;
        }
    }
    return LSD_OK;
}

}  // namespace lsd

using lsd::Ctx;

struct lsd_ctx {
    Ctx* c;
};

#define CTX_OR_FAIL(ctx)                    \
    if (!(ctx) || !(ctx)->c) {              \
        lsd::set_error("null context");     \
        return LSD_E_ARG;                   \
    }                                       \
    Ctx& c = *(ctx)->c;                     \
    if (hipSetDevice(c.device) != hipSuccess) { \
        lsd::set_error("hipSetDevice failed"); \
        return LSD_E_HIP;                   \
    }

extern "C" {

const char* lsd_last_error(void) { return lsd::g_err.c_str(); }

int lsd_config_sizes(const lsd_config* cfg, int* H, int* Th, int* Ph) { return lsd::check_config(cfg, H, Th, Ph); }

int lsd_create(const lsd_config* cfg, int device, lsd_ctx** out) {
    if (!out) {
        lsd::set_error("null output handle");
        return LSD_E_ARG;
    }
    *out = nullptr;
    Ctx* c = nullptr;
    int rc = lsd::create(cfg, device, &c);
    if (rc) {
        std::string keep = lsd::g_err;
        lsd::release(c);
        lsd::set_error(keep);
        return rc;
    }
    *out = new lsd_ctx{c};
    return LSD_OK;
}

void lsd_destroy(lsd_ctx* ctx) {
    if (!ctx) return;
    lsd::release(ctx->c);
    delete ctx;
}

int lsd_sync(lsd_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_set_coll(lsd_ctx* ctx, int field, const double* re, const double* im) {
    CTX_OR_FAIL(ctx);
    if (field < 0 || field >= LSD_NUM_COLL || !re || !im) {
        lsd::set_error("lsd_set_coll: bad field or null pointer");
        return LSD_E_ARG;
    }
    size_t HN = c.coll_elems();
    std::vector<double2> tmp(HN);
    for (size_t q = 0; q < HN; q++) tmp[q] = make_double2(re[q], im[q]);
    LSD_HIP(hipMemcpyAsync(c.d_f[field], tmp.data(), sizeof(double2) * HN, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_set_ic(lsd_ctx* ctx, int which, const double* re, const double* im) {
    CTX_OR_FAIL(ctx);
    if (!c.ic() || which < 0 || which > 1 || !re || !im) {
        lsd::set_error("lsd_set_ic: no conducting inner core, bad field or null pointer");
        return LSD_E_ARG;
    }
    const size_t n = (size_t)c.H * c.Nic;
    std::vector<double2> tmp(n);
    for (size_t q = 0; q < n; q++) tmp[q] = make_double2(re[q], im[q]);
    LSD_HIP(hipMemcpyAsync(c.d_icf[which], tmp.data(), sizeof(double2) * n, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_get_ic(lsd_ctx* ctx, int which, double* re, double* im) {
    CTX_OR_FAIL(ctx);
    if (!c.ic() || which < 0 || which > 1 || !re || !im) {
        lsd::set_error("lsd_get_ic: no conducting inner core, bad field or null pointer");
        return LSD_E_ARG;
    }
    const size_t n = (size_t)c.H * c.Nic;
    std::vector<double2> tmp(n);
    LSD_HIP(hipMemcpyAsync(tmp.data(), c.d_icf[which], sizeof(double2) * n, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    for (size_t q = 0; q < n; q++) {
        re[q] = tmp[q].x;
        im[q] = tmp[q].y;
    }
    return LSD_OK;
}

int lsd_get_rotation(lsd_ctx* ctx, double* out) {
    CTX_OR_FAIL(ctx);
    if (!out) {
        lsd::set_error("lsd_get_rotation: null output");
        return LSD_E_ARG;
    }
    double r[3];
    LSD_HIP(hipMemcpyAsync(r, c.d_rot, sizeof(double) * 3, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 3; k++) out[k] = r[k];
    return LSD_OK;
}

int lsd_set_rotation(lsd_ctx* ctx, double omega) {
    CTX_OR_FAIL(ctx);
    LSD_HIP(hipMemcpyAsync(c.d_rot + lsd::ROT_OMEGA, &omega, sizeof(double), hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_read_state(const char* path, const lsd_config* cfg, double* coll, double* t, double* dt) {
    if (!path || !coll) {
        lsd::set_error("lsd_read_state: null path or output");
        return LSD_E_ARG;
    }
    int rc = lsd::check_config(cfg, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return lsd::state_read(path, *cfg, coll, t, dt);
}

int lsd_write_state(const char* path, const lsd_config* cfg, const double* coll, double t, double dt) {
    if (!path || !coll) {
        lsd::set_error("lsd_write_state: null path or input");
        return LSD_E_ARG;
    }
    int rc = lsd::check_config(cfg, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return lsd::state_write(path, *cfg, coll, t, dt);
}

namespace lsd {

// Host copy of a device array.
static int to_host(const Ctx& c, const double* d, size_t n, std::vector<double>& out) {
    out.resize(n);
    LSD_HIP(hipMemcpyAsync(out.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

// io_ic_extend (program/io.F90:1438-1477): an inner-core field a r^3 + b r^2
// that meets the outer-core field at r_i in value and in the slope of the
// first row of the FD weights dr(1); oc [2][H][N] (Re, Im), ic [2][H][Nic].
static int ic_extend(const Ctx& c, const double* oc, double* ic) {
    std::vector<double> dr1, r, icr;
    int rc = to_host(c, c.d_dr1, (size_t)7 * c.N, dr1);
    if (!rc) rc = to_host(c, c.d_r, (size_t)7 * c.N, r);
    if (!rc) rc = to_host(c, c.d_icr, (size_t)7 * c.Nic, icr);
    if (rc) return rc;
    const int N = c.N, Ni = c.Nic;
    const double ri = r[4 * N], ri2 = r[5 * N], ri3 = r[6 * N];
    for (int part = 0; part < 2; part++)
        for (int h = 0; h < c.H; h++) {
            const double* o = oc + ((size_t)part * c.H + h) * N;
            double* d = ic + ((size_t)part * c.H + h) * Ni;
            const double f = o[0];
            double f1 = 0.0;
            for (int j = 1; j <= 1 + LSD_KL; j++) f1 = f1 + dr1[(j - 1) * 7 + (LSD_KL + 1 - j)] * o[j - 1];
            const double a = (f1 * ri - 2.0 * f) / ri3;
            const double b = (f - a * ri3) / ri2;
            for (int n = 0; n < Ni; n++) d[n] = a * icr[6 * Ni + n] + b * icr[5 * Ni + n];
        }
    return LSD_OK;
}

static StateIc state_ic(const Ctx& c) {
    StateIc ic;
    ic.Nic = c.ic() ? c.Nic : 1;
    if (c.ic()) {
        std::vector<double> icr;
        if (!to_host(c, c.d_icr, (size_t)7 * c.Nic, icr)) ic.ric.assign(icr.begin() + 4 * c.Nic, icr.begin() + 5 * c.Nic);
    }
    return ic;
}

}  // namespace lsd

int lsd_load_state(lsd_ctx* ctx, const char* path) {
    CTX_OR_FAIL(ctx);
    if (c.it != 0) {
        lsd::set_error("lsd_load_state: only between timesteps (tim_it = 0)");
        return LSD_E_ARG;
    }
    const size_t HN = c.coll_elems();
    std::vector<double> buf(12 * HN);
    double t = 0.0, dt = 0.0;
    lsd::StateIc ic = lsd::state_ic(c);
    int rc = lsd::state_read(path, c.cfg, buf.data(), &t, &dt, &ic);
    if (rc) return rc;
    for (int k = 0; k < 6; k++) {
        rc = lsd_set_coll(ctx, k, buf.data() + (size_t)(2 * k) * HN, buf.data() + (size_t)(2 * k + 1) * HN);
        if (rc) return rc;
    }
    if (c.ic()) {
        // the file's inner-core field, or the outer core's extended into the
        // inner core (program/io.F90:440-447)
        const size_t HI = (size_t)c.H * c.Nic;
        if (!ic.loaded) {
            ic.f.assign(4 * HI, 0.0);
            for (int k = 0; k < 2; k++) {
                rc = lsd::ic_extend(c, buf.data() + (size_t)(2 * (4 + k)) * HN, ic.f.data() + (size_t)(2 * k) * HI);
                if (rc) return rc;
            }
        }
        for (int k = 0; k < 2; k++) {
            rc = lsd_set_ic(ctx, k, ic.f.data() + (size_t)(2 * k) * HI, ic.f.data() + (size_t)(2 * k + 1) * HI);
            if (rc) return rc;
        }
    }
    // rot_omega from icOmega (program/io.F90:430-431)
    if (c.cfg.rot || c.ic()) {
        rc = lsd_set_rotation(ctx, ic.omega);
        if (rc) return rc;
    }
    // program/io.F90:394-430: tim_t from the file; with d_timestep < 0 also
    // tim_dt, and the operators follow it
    c.t = t;
    if (c.adaptive && dt > 0.0 && dt != c.dt) {
        c.dt = dt;
        rc = lsd::build_operators(c);
        if (rc) return rc;
    }
    return LSD_OK;
}

int lsd_save_state(lsd_ctx* ctx, const char* path) {
    CTX_OR_FAIL(ctx);
    const size_t HN = c.coll_elems();
    std::vector<double> buf(12 * HN);
    for (int k = 0; k < 6; k++) {
        int rc = lsd_get_coll(ctx, k, buf.data() + (size_t)(2 * k) * HN, buf.data() + (size_t)(2 * k + 1) * HN);
        if (rc) return rc;
    }
    lsd::StateIc ic = lsd::state_ic(c);
    if (c.ic()) {
        const size_t HI = (size_t)c.H * c.Nic;
        ic.f.assign(4 * HI, 0.0);
        for (int k = 0; k < 2; k++) {
            int rc = lsd_get_ic(ctx, k, ic.f.data() + (size_t)(2 * k) * HI, ic.f.data() + (size_t)(2 * k + 1) * HI);
            if (rc) return rc;
        }
    }
    double rot[3] = {0.0, 0.0, 0.0};
    int rc = lsd_get_rotation(ctx, rot);
    if (rc) return rc;
    ic.omega = rot[0];
    return lsd::state_write(path, c.cfg, buf.data(), c.t, c.dt, &ic);
}

int lsd_get_coll(lsd_ctx* ctx, int field, double* re, double* im) {
    CTX_OR_FAIL(ctx);
    if (field < 0 || field >= LSD_NUM_COLL || !re || !im) {
        lsd::set_error("lsd_get_coll: bad field or null pointer");
        return LSD_E_ARG;
    }
    size_t HN = c.coll_elems();
    std::vector<double2> tmp(HN);
    LSD_HIP(hipMemcpyAsync(tmp.data(), c.d_f[field], sizeof(double2) * HN, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    for (size_t q = 0; q < HN; q++) {
        re[q] = tmp[q].x;
        im[q] = tmp[q].y;
    }
    return LSD_OK;
}

int lsd_get_phys(lsd_ctx* ctx, int field, double* p) {
    CTX_OR_FAIL(ctx);
    if (field < 0 || field >= LSD_NUM_PHYS || !p) {
        lsd::set_error("lsd_get_phys: bad field or null pointer");
        return LSD_E_ARG;
    }
    LSD_HIP(hipMemcpyAsync(p, lsd::phys(c, field), sizeof(double) * c.phys_elems(), hipMemcpyDeviceToHost,
                           c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_get_table(lsd_ctx* ctx, int table, double* out) {
    CTX_OR_FAIL(ctx);
    if (!out) {
        lsd::set_error("null output");
        return LSD_E_ARG;
    }
    size_t HT = (size_t)c.H * c.Th2;
    if (table == LSD_T_GP || table == LSD_T_GW) {
        LSD_HIP(hipMemcpyAsync(out, table == LSD_T_GP ? c.d_gp : c.d_gw, sizeof(double) * c.Th,
                               hipMemcpyDeviceToHost, c.stream));
    } else if (table >= LSD_T_P && table <= LSD_T_LLSIN1MP) {
        LSD_HIP(hipMemcpyAsync(out, c.d_tab + (size_t)(table - LSD_T_P) * HT, sizeof(double) * HT,
                               hipMemcpyDeviceToHost, c.stream));
    } else if (table >= LSD_T_GWP && table <= LSD_T_GWLLSIN1MP) {
        std::vector<double> tmp(HT);
        LSD_HIP(hipMemcpyAsync(tmp.data(), c.d_atab + (size_t)(table - LSD_T_GWP) * HT, sizeof(double) * HT,
                               hipMemcpyDeviceToHost, c.stream));
        LSD_HIP(hipStreamSynchronize(c.stream));
        for (auto& s : c.slots)
            for (int k = 0; k < s.len; k++) {
                int ap = s.h0 + ((k & 1) ? (s.len + 1) / 2 + (k >> 1) : (k >> 1));
                for (int j = 0; j < c.Th2; j++) out[(size_t)(s.h0 + k) * c.Th2 + j] = tmp[(size_t)j * c.H + ap];
            }
    } else {
        lsd::set_error("lsd_get_table: unknown table");
        return LSD_E_ARG;
    }
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_get_mesh(lsd_ctx* ctx, int which, double* out) {
    CTX_OR_FAIL(ctx);
    if (!out) {
        lsd::set_error("null output");
        return LSD_E_ARG;
    }
    const double* src = nullptr;
    size_t n = 7 * (size_t)c.N;
    switch (which) {
    case LSD_M_R: src = c.d_r + 4 * c.N; n = c.N; break;
    case LSD_M_DR1: src = c.d_dr1; break;
    case LSD_M_DR2: src = c.d_dr2; break;
    case LSD_M_RADLAP: src = c.d_radlap; break;
    case LSD_M_R1DR: src = c.d_r1dr; break;
    case LSD_M_INTR2DR: src = c.d_intr2dr; n = c.N; break;
    default: lsd::set_error("lsd_get_mesh: unknown array"); return LSD_E_ARG;
    }
    LSD_HIP(hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_transform(lsd_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return lsd::transform(c);
}
int lsd_nonlinear(lsd_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return lsd::nonlinear(c);
}
// The phase calls keep the context's tim_it in step with the caller's:
// the predictor opens a step, and each corrector pass applies the
// tim_check_cgce rule to its tim_dterr (the caller's own check decides
// whether the run stops), so lsd_set_dt can refuse a step in progress.
int lsd_predictor(lsd_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    int rc = lsd::launch_predictor(c);
    if (rc) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    c.it = 1;
    return LSD_OK;
}
int lsd_corrector(lsd_ctx* ctx, double* dterr) {
    CTX_OR_FAIL(ctx);
    double d = 0.0;
    int rc = lsd::launch_corrector(c, &d);
    if (rc) return rc;
    if (dterr) *dterr = d;
    if (c.it != 0) {
        c.dterr += d;
        (void)lsd::check_cgce(c);
    }
    return LSD_OK;
}

int lsd_energy(lsd_ctx* ctx, int which, double* E, double* El, double* Em) {
    CTX_OR_FAIL(ctx);
    if ((which != 0 && which != 1) || !E) {
        lsd::set_error("lsd_energy: which must be 0 (velocity) or 1 (magnetic), E non-null");
        return LSD_E_ARG;
    }
    if (which == 1 && !c.cfg.mag) {
        lsd::set_error("lsd_energy: the magnetic field is not time-stepped");
        return LSD_E_ARG;
    }
    return lsd::energy(c, which, E, El, Em);
}

int lsd_energy_noalloc(lsd_ctx* ctx, int which, double* E, double* El, double* Em) {
    CTX_OR_FAIL(ctx);
    if ((which != 0 && which != 1) || !E) {
        lsd::set_error("lsd_energy_noalloc: which must be 0 (velocity) or 1 (magnetic), E non-null");
        return LSD_E_ARG;
    }
    if (which == 1 && !c.cfg.mag) {
        lsd::set_error("lsd_energy_noalloc: the magnetic field is not time-stepped");
        return LSD_E_ARG;
    }
    return lsd::energy_noalloc(c, which, E, El, Em);
}

int lsd_step_many(lsd_ctx** ctxs, int n, int nsteps, lsd_step_stats* stats) {
    if (!ctxs || n < 1) {
        lsd::set_error("lsd_step_many: no contexts");
        return LSD_E_ARG;
    }
    std::vector<Ctx*> cs;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !ctxs[i]->c) {
            lsd::set_error("lsd_step_many: null context");
            return LSD_E_ARG;
        }
        for (Ctx* o : cs)
            if (o == ctxs[i]->c) {
                lsd::set_error("lsd_step_many: a context appears twice");
                return LSD_E_ARG;
            }
        cs.push_back(ctxs[i]->c);
    }
    return lsd::step_many(cs, nsteps, stats);
}

int lsd_set_graph(lsd_ctx* ctx, int on) {
    CTX_OR_FAIL(ctx);
    LSD_HIP(hipStreamSynchronize(c.stream));
    c.use_graph = on != 0;
    return LSD_OK;
}

int lsd_step(lsd_ctx* ctx, int nsteps, lsd_step_stats* stats) {
    CTX_OR_FAIL(ctx);
    int rc = lsd::step(c, nsteps, stats);
    if (rc) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_profile_begin(lsd_ctx* ctx, int kernel_class) {
    CTX_OR_FAIL(ctx);
    if (kernel_class < 0 || kernel_class >= LSD_K_COUNT) {
        lsd::set_error("lsd_profile_begin: unknown kernel class");
        return LSD_E_ARG;
    }
    LSD_HIP(hipStreamSynchronize(c.stream));
    c.prof_class = kernel_class;
    c.prof_used = 0;
    c.prof_work = 0.0;
    c.prof_acc_ms = c.prof_acc_work = 0.0;
    c.prof_acc_launches = 0;
    return LSD_OK;
}

int lsd_profile_end(lsd_ctx* ctx, double* total_ms, int* launches, double* work) {
    CTX_OR_FAIL(ctx);
    LSD_HIP(hipStreamSynchronize(c.stream));
    double tot = 0.0;
    for (int k = 0; k + 1 < c.prof_used; k += 2) {
        float ms = 0.f;
        LSD_HIP(hipEventElapsedTime(&ms, c.prof_events[k], c.prof_events[k + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot + c.prof_acc_ms;
    if (launches) *launches = c.prof_used / 2 + c.prof_acc_launches;
    if (work) *work = c.prof_work + c.prof_acc_work;
    c.prof_class = -1;
    c.prof_used = 0;
    c.prof_work = 0.0;
    c.prof_acc_ms = c.prof_acc_work = 0.0;
    c.prof_acc_launches = 0;
    return LSD_OK;
}

int lsd_set_variant(lsd_ctx* ctx, int which, int value) {
    CTX_OR_FAIL(ctx);
    if (c.it != 0) {
        lsd::set_error("lsd_set_variant: only between timesteps (tim_it = 0)");
        return LSD_E_ARG;
    }
    if (!lsd::variant_ok(c, which, value)) {
        lsd::set_error("lsd_set_variant: variant " + std::to_string(which) + " = " + std::to_string(value) +
                       " does not apply to this context");
        return LSD_E_ARG;
    }
    c.var[which] = value;
    lsd::drop_graphs(c);
    lsd::select_solve_form(c);
    return LSD_OK;
}

int lsd_get_variant(lsd_ctx* ctx, int which, int* value) {
    CTX_OR_FAIL(ctx);
    if (which < 0 || which >= LSD_V_COUNT || !value) {
        lsd::set_error("lsd_get_variant: bad variant or null output");
        return LSD_E_ARG;
    }
    *value = c.var[which];
    return LSD_OK;
}

int lsd_tune(lsd_ctx* ctx, int reps, const char* cache_path, double* times_us) {
    CTX_OR_FAIL(ctx);
    return lsd::tune(c, reps, cache_path, times_us);
}

int lsd_get_dt(lsd_ctx* ctx, double* dt, double* cfl_dt, double* corr_dt) {
    CTX_OR_FAIL(ctx);
    if (dt) *dt = c.dt;
    if (cfl_dt) *cfl_dt = c.cfl_dt;
    if (corr_dt) *corr_dt = c.corr_dt;
    return LSD_OK;
}

int lsd_set_dt(lsd_ctx* ctx, double dt) {
    CTX_OR_FAIL(ctx);
    if (!(dt > 0.0)) {
        lsd::set_error("dt must be > 0");
        return LSD_E_ARG;
    }
    if (c.it != 0) {
        lsd::set_error("lsd_set_dt: only between timesteps (tim_it = 0)");
        return LSD_E_ARG;
    }
    c.dt = dt;
    if (!c.adaptive) c.cfg.dt = dt;
    return lsd::build_operators(c);
}

// ---------------------------------------------------- drop-in transforms

static int upload_spec(Ctx& c, const double* re, const double* im, double2* dst) {
    size_t HN = c.coll_elems();
    double* tmp = reinterpret_cast<double*>(lsd::nlphys(c, 0));  // staging, >= 2*H*N doubles
    LSD_HIP(hipMemcpyAsync(tmp, re, sizeof(double) * HN, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipMemcpyAsync(tmp + HN, im, sizeof(double) * HN, hipMemcpyHostToDevice, c.stream));
    return lsd::launch_layout_spec2coll(c, tmp, tmp + HN, dst);
}

static int download_spec(Ctx& c, const double2* src, double* re, double* im) {
    size_t HN = c.coll_elems();
    double* tmp = reinterpret_cast<double*>(lsd::nlphys(c, 0));
    int rc = lsd::launch_layout_coll2spec(c, src, tmp, tmp + HN);
    if (rc) return rc;
    LSD_HIP(hipMemcpyAsync(re, tmp, sizeof(double) * HN, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipMemcpyAsync(im, tmp + HN, sizeof(double) * HN, hipMemcpyDeviceToHost, c.stream));
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

static int download_phys(Ctx& c, int k, double* p) {
    LSD_HIP(hipMemcpyAsync(p, lsd::phys(c, k), sizeof(double) * c.phys_elems(), hipMemcpyDeviceToHost, c.stream));
    return LSD_OK;
}

// Synthesis -> analysis round trip at the Fourier level in the context's
// precision (FP64, FP32 or BF16x3 Legendre kernels).
int lsd_sht_roundtrip(lsd_ctx* ctx, int kind, const double* in, double* out) {
    CTX_OR_FAIL(ctx);
    if ((kind != 0 && kind != 1) || !in || !out) {
        lsd::set_error("lsd_sht_roundtrip: kind must be 0 (scalar) or 1 (vector), non-null arrays");
        return LSD_E_ARG;
    }
    const bool f = c.mixed;
    const int nf = kind ? 3 : 1;
    const size_t HN = c.coll_elems();
    double2* stage = c.d_f[LSD_N_VEL_TOR];  // staging: the N fields are recomputed by the next evaluation
    std::vector<double2> tmp(HN);
    for (int k = 0; k < nf; k++) {
        const double* re = in + (size_t)(2 * k) * HN;
        const double* im = in + (size_t)(2 * k + 1) * HN;
        for (size_t q = 0; q < HN; q++) tmp[q] = make_double2(re[q], im[q]);
        LSD_HIP(hipMemcpyAsync(stage, tmp.data(), sizeof(double2) * HN, hipMemcpyHostToDevice, c.stream));
        int rc = lsd::launch_cast_coll(c, stage, lsd::spec_in_p(c, k, f), f, false);
        if (rc) return rc;
        LSD_HIP(hipStreamSynchronize(c.stream));
    }
    int rc;
    if (kind == 0) {
        if ((rc = lsd::launch_synthesis_p(c, lsd::SYN_SCALAR, lsd::spec_in_p(c, 0, f), nullptr,
                                          lsd::fourier_p(c, 0, f), nullptr, f)))
            return rc;
        if ((rc = lsd::launch_fold(c, lsd::fourier_p(c, 0, f), lsd::folded_p(c, 0, f), f))) return rc;
        if ((rc = lsd::launch_analysis_p(c, lsd::ANA_SCALAR, lsd::folded_p(c, 0, f), nullptr, nullptr,
                                         lsd::spec_out_p(c, 0, f), nullptr, nullptr, f)))
            return rc;
    } else {
        if ((rc = lsd::launch_synthesis_p(c, lsd::SYN_SCALAR, lsd::spec_in_p(c, 0, f), nullptr,
                                          lsd::fourier_p(c, 0, f), nullptr, f)) ||
            (rc = lsd::launch_synthesis_p(c, lsd::SYN_VECTOR, lsd::spec_in_p(c, 1, f), lsd::spec_in_p(c, 2, f),
                                          lsd::fourier_p(c, 1, f), lsd::fourier_p(c, 2, f), f)))
            return rc;
        for (int k = 0; k < 3; k++)
            if ((rc = lsd::launch_fold(c, lsd::fourier_p(c, k, f), lsd::folded_p(c, k, f), f))) return rc;
        if ((rc = lsd::launch_analysis_p(c, lsd::ANA_VECTOR, lsd::folded_p(c, 0, f), lsd::folded_p(c, 1, f),
                                         lsd::folded_p(c, 2, f), lsd::spec_out_p(c, 0, f), lsd::spec_out_p(c, 1, f),
                                         lsd::spec_out_p(c, 2, f), f)))
            return rc;
    }
    for (int k = 0; k < nf; k++) {
        if ((rc = lsd::launch_cast_coll(c, lsd::spec_out_p(c, k, f), stage, false, f))) return rc;
        LSD_HIP(hipMemcpyAsync(tmp.data(), stage, sizeof(double2) * HN, hipMemcpyDeviceToHost, c.stream));
        LSD_HIP(hipStreamSynchronize(c.stream));
        double* re = out + (size_t)(2 * k) * HN;
        double* im = out + (size_t)(2 * k + 1) * HN;
        for (size_t q = 0; q < HN; q++) {
            re[q] = tmp[q].x;
            im[q] = tmp[q].y;
        }
    }
    return LSD_OK;
}

int lsd_tra_spec2phys(lsd_ctx* ctx, const double* s_re, const double* s_im, double* p) {
    CTX_OR_FAIL(ctx);
    int rc;
    double2* in = lsd::spec_in(c, lsd::SI_DC);
    if ((rc = upload_spec(c, s_re, s_im, in))) return rc;
    if ((rc = lsd::launch_synthesis(c, lsd::SYN_SCALAR, in, nullptr, lsd::fourier(c, 0), nullptr))) return rc;
    if ((rc = lsd::launch_fft_c2r(c, lsd::fourier(c, 0), lsd::phys(c, 0), 1))) return rc;
    if ((rc = download_phys(c, 0, p))) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_tra_grad(lsd_ctx* ctx, const double* s_re, const double* s_im, double* t, double* p) {
    CTX_OR_FAIL(ctx);
    int rc;
    double2* in = lsd::spec_in(c, lsd::SI_DC);
    if ((rc = upload_spec(c, s_re, s_im, in))) return rc;
    if ((rc = lsd::launch_synthesis(c, lsd::SYN_GRAD, in, nullptr, lsd::fourier(c, 0), lsd::fourier(c, 1))))
        return rc;
    if ((rc = lsd::launch_fft_c2r(c, lsd::fourier(c, 0), lsd::phys(c, 0), 2))) return rc;
    if ((rc = download_phys(c, 0, t))) return rc;
    if ((rc = download_phys(c, 1, p))) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_tra_qst2rtp(lsd_ctx* ctx, const double* q_re, const double* q_im, const double* s_re, const double* s_im,
                    const double* t_re, const double* t_im, double* r, double* th, double* ph) {
    CTX_OR_FAIL(ctx);
    int rc;
    double2* q = lsd::spec_in(c, lsd::SI_UQ);
    double2* s = lsd::spec_in(c, lsd::SI_US);
    double2* t = lsd::spec_in(c, lsd::SI_UT);
    if ((rc = upload_spec(c, q_re, q_im, q)) || (rc = upload_spec(c, s_re, s_im, s)) ||
        (rc = upload_spec(c, t_re, t_im, t)))
        return rc;
    if ((rc = lsd::launch_synthesis(c, lsd::SYN_SCALAR, q, nullptr, lsd::fourier(c, 0), nullptr))) return rc;
    if ((rc = lsd::launch_synthesis(c, lsd::SYN_VECTOR, s, t, lsd::fourier(c, 1), lsd::fourier(c, 2)))) return rc;
    if ((rc = lsd::launch_fft_c2r(c, lsd::fourier(c, 0), lsd::phys(c, 0), 3))) return rc;
    if ((rc = download_phys(c, 0, r)) || (rc = download_phys(c, 1, th)) || (rc = download_phys(c, 2, ph))) return rc;
    LSD_HIP(hipStreamSynchronize(c.stream));
    return LSD_OK;
}

int lsd_tra_phys2spec(lsd_ctx* ctx, const double* p, double* s_re, double* s_im) {
    CTX_OR_FAIL(ctx);
    int rc;
    double* x = lsd::phys(c, 0);
    LSD_HIP(hipMemcpyAsync(x, p, sizeof(double) * c.phys_elems(), hipMemcpyHostToDevice, c.stream));
    if ((rc = lsd::launch_fft_r2c_fold(c, x, lsd::folded(c, 0), 1))) return rc;
    double2* o = lsd::spec_out(c, lsd::SO_C);
    if ((rc = lsd::launch_analysis(c, lsd::ANA_SCALAR, lsd::folded(c, 0), nullptr, nullptr, o, nullptr, nullptr)))
        return rc;
    return download_spec(c, o, s_re, s_im);
}

int lsd_tra_rtp2qst(lsd_ctx* ctx, const double* r, const double* th, const double* ph, double* q_re, double* q_im,
                    double* s_re, double* s_im, double* t_re, double* t_im) {
    CTX_OR_FAIL(ctx);
    int rc;
    size_t np = c.phys_elems();
    LSD_HIP(hipMemcpyAsync(lsd::phys(c, 0), r, sizeof(double) * np, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipMemcpyAsync(lsd::phys(c, 1), th, sizeof(double) * np, hipMemcpyHostToDevice, c.stream));
    LSD_HIP(hipMemcpyAsync(lsd::phys(c, 2), ph, sizeof(double) * np, hipMemcpyHostToDevice, c.stream));
    if ((rc = lsd::launch_fft_r2c_fold(c, lsd::phys(c, 0), lsd::folded(c, 0), 3))) return rc;
    double2* o0 = lsd::spec_out(c, lsd::SO_UQ);
    double2* o1 = lsd::spec_out(c, lsd::SO_US);
    double2* o2 = lsd::spec_out(c, lsd::SO_UT);
    if ((rc = lsd::launch_analysis(c, lsd::ANA_VECTOR, lsd::folded(c, 0), lsd::folded(c, 1), lsd::folded(c, 2), o0,
                                   o1, o2)))
        return rc;
    if ((rc = download_spec(c, o0, q_re, q_im)) || (rc = download_spec(c, o1, s_re, s_im)) ||
        (rc = download_spec(c, o2, t_re, t_im)))
        return rc;
    return LSD_OK;
}

}  // extern "C"
