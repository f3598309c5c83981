// C entry points: argument validation, plan/model construction and the
// orchestration of the per-stage kernels.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "plan.hpp"
#include "split.hpp"

namespace aeo {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
const char* get_error() { return g_last_error.c_str(); }

int fft_factor(int n, int* radices) {
    int k = 0;
    while (n % 4 == 0 && k < kMaxRadices) { radices[k++] = 4; n /= 4; }
    while (n % 2 == 0 && k < kMaxRadices) { radices[k++] = 2; n /= 2; }
    while (n % 3 == 0 && k < kMaxRadices) { radices[k++] = 3; n /= 3; }
    while (n % 5 == 0 && k < kMaxRadices) { radices[k++] = 5; n /= 5; }
    for (int p = 7; n > 1 && k < kMaxRadices; p += 2) {
        while (n % p == 0 && k < kMaxRadices) { radices[k++] = p; n /= p; }
    }
    return n == 1 ? k : -1;
}

template <class T>
static int dev_alloc(T** ptr, size_t count) {
    *ptr = nullptr;
    if (count == 0) return AEO_OK;
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                  " bytes failed: " + hipGetErrorString(e));
        *ptr = nullptr;
        return AEO_ERR_HIP;
    }
    return AEO_OK;
}

template <class T>
static int upload(T** ptr, const T* host, size_t count) {
    int rc = dev_alloc(ptr, count);
    if (rc) return rc;
    AEO_CHECK_HIP(hipMemcpy(*ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
    return AEO_OK;
}

static void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

static void free_plan(aeo_plan_s* p) {
    if (!p) return;
    for (void* q : {(void*)p->d_Ps, (void*)p->d_Pa, (void*)p->d_offs, (void*)p->d_w_rows,
                    (void*)p->d_coslat, (void*)p->d_tw, (void*)p->d_b1pa, (void*)p->d_E,
                    (void*)p->d_O,
                    (void*)p->d_S, (void*)p->d_D, (void*)p->d_Yk, (void*)p->d_Yp, (void*)p->d_w,
                    (void*)p->d_pcoslat, (void*)p->d_gw, (void*)p->d_part, (void*)p->d_Cp,
                    (void*)p->d_Cm, (void*)p->d_spec, (void*)p->d_Psl, (void*)p->d_offs_ss,
                    (void*)p->d_esyn, (void*)p->d_Pasl, (void*)p->d_offs_sa, (void*)p->d_eana,
                    (void*)p->d_offs_ea, (void*)p->d_xn, (void*)p->d_rec, (void*)p->d_CpT,
                    (void*)p->d_CmT, (void*)p->d_Ysl, (void*)p->d_eYs, (void*)p->d_Yasl,
                    (void*)p->d_eYa})
        dev_free(q);
    delete p;
}

static void free_prof(ProfRec& pr) {
    if (pr.ev) {
        for (int i = 0; i < 2 * pr.cap; ++i) (void)hipEventDestroy(pr.ev[i]);
        delete[] pr.ev;
    }
    pr = ProfRec();
}

static void free_model(aeo_model_s* m) {
    if (!m) return;
    free_prof(m->prof);
    for (void* q : {(void*)m->d_nulap, (void*)m->d_lap, (void*)m->d_nu4lap2,
                    (void*)m->d_lap_phi_s, (void*)m->d_nu4_lap2_phi_s, (void*)m->d_E,
                    (void*)m->d_O, (void*)m->d_E1, (void*)m->d_O1, (void*)m->d_S,
                    (void*)m->d_D, (void*)m->d_gw,
                    (void*)m->d_part, (void*)m->d_ystage, (void*)m->d_ybuf, (void*)m->d_ksum,
                    (void*)m->d_fint, (void*)m->d_aE, (void*)m->d_aO, (void*)m->d_aS,
                    (void*)m->d_aD, (void*)m->d_asum, (void*)m->d_abuf, (void*)m->d_kbuf,
                    (void*)m->d_gT, (void*)m->d_gQ, (void*)m->d_gQe, (void*)m->d_gkoff})
        dev_free(q);
    delete m;
}

// C+/C- coupling coefficients of sin(theta) d/dtheta
// (num/spectral_operators.py:135-176), row-major (L+1)^2.
static void coupling_tables(int L, std::vector<double>& Cp, std::vector<double>& Cm) {
    const int n = L + 1;
    Cp.assign((size_t)n * n, 0.0);
    Cm.assign((size_t)n * n, 0.0);
    for (int l = 0; l <= L; ++l) {
        for (int m = 0; m <= l; ++m) {
            const double Lb = l, Mb = m;
            const double np_ = std::max((Lb + 1.0) * (Lb + 1.0) - Mb * Mb, 0.0);
            const double dp = std::max((2.0 * Lb + 1.0) * (2.0 * Lb + 3.0), 1.0);
            const double nm = std::max(Lb * Lb - Mb * Mb, 0.0);
            const double dm = std::max((2.0 * Lb - 1.0) * (2.0 * Lb + 1.0), 1.0);
            Cp[(size_t)l * n + m] = (l == L) ? 0.0 : Lb * std::sqrt(np_ / dp);
            Cm[(size_t)l * n + m] = (l == 0) ? 0.0 : -(Lb + 1.0) * std::sqrt(nm / dm);
        }
    }
}

static int common_tables(aeo_plan_s* p) {
    std::vector<double> Cp, Cm;
    coupling_tables(p->L, Cp, Cm);
    int rc = upload(&p->d_Cp, Cp.data(), Cp.size());
    if (rc) return rc;
    rc = upload(&p->d_Cm, Cm.data(), Cm.size());
    if (rc) return rc;
    // column-major copies: the stage kernels read one column m per workgroup
    const int n = p->L + 1;
    std::vector<double> CpT((size_t)n * n), CmT((size_t)n * n);
    for (int l = 0; l < n; ++l)
        for (int m = 0; m < n; ++m) {
            CpT[(size_t)m * n + l] = Cp[(size_t)l * n + m];
            CmT[(size_t)m * n + l] = Cm[(size_t)l * n + m];
        }
    if ((rc = upload(&p->d_CpT, CpT.data(), CpT.size()))) return rc;
    if ((rc = upload(&p->d_CmT, CmT.data(), CmT.size()))) return rc;
    const int nw = std::max(p->max_fields, 4);
    return dev_alloc(&p->d_spec, (size_t)nw * (p->L + 1) * (p->L + 1));
}

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace aeo

using namespace aeo;

extern "C" {

int aeo_version(void) { return AEO_ABI_VERSION; }

const char* aeo_last_error(void) { return get_error(); }

int aeo_plan_create_gauss(int nlat, int nlon, int l_max, const double* x_nodes,
                          const double* w_rows, const double* coslat_rows, int max_fields,
                          aeo_plan_t* out) {
    AEO_REQUIRE(out != nullptr, AEO_ERR_INVALID, "aeo_plan_create_gauss: out is NULL");
    *out = nullptr;
    AEO_REQUIRE(x_nodes && w_rows && coslat_rows, AEO_ERR_INVALID,
                "aeo_plan_create_gauss: x_nodes, w_rows and coslat_rows must be non-NULL");
    AEO_REQUIRE(nlat >= 2 && nlon >= 4, AEO_ERR_INVALID,
                "aeo_plan_create_gauss: needs nlat >= 2 and nlon >= 4");
    AEO_REQUIRE(l_max >= 0, AEO_ERR_INVALID, "aeo_plan_create_gauss: l_max must be >= 0");
    AEO_REQUIRE(max_fields >= 1, AEO_ERR_INVALID, "aeo_plan_create_gauss: max_fields must be >= 1");
    AEO_REQUIRE(nlon <= kChunkElems, AEO_ERR_UNSUPPORTED,
                "aeo_plan_create_gauss: nlon > 2048 is not supported by the LDS FFT");
    int radices[kMaxRadices];
    const int nst = fft_factor(nlon, radices);
    AEO_REQUIRE(nst > 0, AEO_ERR_UNSUPPORTED,
                "aeo_plan_create_gauss: nlon has too many prime factors for the LDS FFT");
    for (int j = 0; j < nlat; ++j) {
        const int js = nlat - 1 - j;
        AEO_REQUIRE(std::fabs(x_nodes[j] + x_nodes[js]) <= 1e-15 &&
                        std::fabs(w_rows[j] - w_rows[js]) <= 1e-15 * std::fabs(w_rows[j]) &&
                        (j == 0 || x_nodes[j] < x_nodes[j - 1]),
                    AEO_ERR_INVALID,
                    "aeo_plan_create_gauss: nodes must be descending and symmetric about the "
                    "equator (x[j] = -x[nlat-1-j], equal weights)");
    }
    aeo_plan_s* p = new (std::nothrow) aeo_plan_s();
    AEO_REQUIRE(p != nullptr, AEO_ERR_HIP, "aeo_plan_create_gauss: out of host memory");
    p->kind = kGauss;
    p->L = l_max;
    p->B = n_basis(l_max);
    p->nlat = nlat;
    p->nlon = nlon;
    p->nh = (nlat + 1) / 2;
    p->max_fields = max_fields;
    const int L = l_max, nh = p->nh;
    std::vector<long long> offs(L + 2, 0);
    for (int m = 0; m <= L; ++m) offs[m + 1] = offs[m] + (long long)nh * (L - m + 1);
    // tw[t] = e^{-2 pi i t / nlon}; for nlon divisible by 4 the first quarter is
    // evaluated and the rest are its exact quarter-turn rotations, so kernels
    // that keep only the quarter table (twiddle_q) get bitwise the same values
    std::vector<double2> tw(nlon);
    const int q4 = (nlon % 4 == 0) ? nlon / 4 : nlon;
    for (int t = 0; t < q4; ++t) {
        const double ang = 2.0 * M_PI * (double)t / (double)nlon;
        tw[t] = make_double2(std::cos(ang), -std::sin(ang));
    }
    for (int t = q4; t < nlon; ++t) {
        const double2 w = tw[t % q4];
        const int q = t / q4;
        double x = (q & 1) ? w.y : w.x, y = (q & 1) ? -w.x : w.y;
        if (q & 2) {
            x = -x;
            y = -y;
        }
        tw[t] = make_double2(x, y);
    }
    // the per-row Jacobian kernel's stage-major copy (hemi_tw) after it
    if ((nlon & (nlon - 1)) == 0 && nlon >= 64 && nlon <= 512) {
        int lgn = 0;
        while ((1 << lgn) < nlon) ++lgn;
        const HemiTw ht = hemi_tw(lgn);
        for (int b = 0; b < ht.n; ++b)
            for (int r = 1; r < (1 << ht.lgr[b]); ++r)
                for (int k = 0; k < (1 << ht.lgns[b]); ++k) {
                    const double2 w = tw[(r * k) << (lgn - ht.lgns[b] - ht.lgr[b])];
                    tw.push_back(w);
                }
    }
    std::vector<double> xn(x_nodes, x_nodes + nh);
    // Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m}), stored m-major as
    // (a_lm, a_lm b_lm) for the one-fma step of rec_step
    std::vector<double2> rec((size_t)(L + 1) * (L + 1), make_double2(0.0, 0.0));
    for (int m = 0; m <= L; ++m)
        for (int l = m + 2; l <= L; ++l) {
            const double ll = (double)l * l, mm = (double)m * m, l1 = l - 1.0;
            const double a = std::sqrt((4.0 * ll - 1.0) / (ll - mm));
            const double b = std::sqrt((l1 * l1 - mm) / (4.0 * l1 * l1 - 1.0));
            rec[(size_t)m * (L + 1) + l] = make_double2(a, a * b);
        }
    int rc = AEO_OK;
    const size_t plane = (size_t)nh * (L + 1);
    const int nwork = ((std::max(max_fields, 12) + 3) / 4) * 4;
    do {
        if ((rc = upload(&p->d_offs, offs.data(), offs.size()))) break;
        if ((rc = dev_alloc(&p->d_Ps, (size_t)offs[L + 1]))) break;
        if ((rc = dev_alloc(&p->d_Pa, (size_t)offs[L + 1]))) break;
        if ((rc = upload(&p->d_w_rows, w_rows, (size_t)nlat))) break;
        if ((rc = upload(&p->d_coslat, coslat_rows, (size_t)nlat))) break;
        if ((rc = upload(&p->d_tw, tw.data(), tw.size()))) break;
        if ((rc = dev_alloc(&p->d_E, plane * nwork))) break;
        if ((rc = dev_alloc(&p->d_O, plane * nwork))) break;
        if ((rc = dev_alloc(&p->d_S, plane * nwork))) break;
        if ((rc = dev_alloc(&p->d_D, plane * nwork))) break;
        if ((rc = upload(&p->d_xn, xn.data(), xn.size()))) break;
        if ((rc = upload(&p->d_rec, rec.data(), rec.size()))) break;
        if ((rc = common_tables(p))) break;
        p->fft.n = nlon;
        p->fft.nstages = nst;
        for (int i = 0; i < nst; ++i) p->fft.radix[i] = radices[i];
        p->fft.tw = p->d_tw;
        p->fft.lgn = -1;
        if ((nlon & (nlon - 1)) == 0) {
            int lg = 0;
            while ((1 << lg) < nlon) ++lg;
            p->fft.lgn = lg;
        }
        if ((rc = gauss_build_tables(p, nullptr))) break;
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            set_error(std::string("aeo_plan_create_gauss: ") + hipGetErrorString(e));
            rc = AEO_ERR_HIP;
        }
    } while (0);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return AEO_OK;
}

int aeo_plan_create_points(int npoints, int l_max, const double* lat, const double* lon,
                           const double* weights, const double* coslat, int max_fields,
                           aeo_plan_t* out) {
    AEO_REQUIRE(out != nullptr, AEO_ERR_INVALID, "aeo_plan_create_points: out is NULL");
    *out = nullptr;
    AEO_REQUIRE(lat && lon && coslat, AEO_ERR_INVALID,
                "aeo_plan_create_points: lat, lon and coslat must be non-NULL");
    AEO_REQUIRE(npoints >= 1 && l_max >= 0 && max_fields >= 1, AEO_ERR_INVALID,
                "aeo_plan_create_points: needs npoints >= 1, l_max >= 0, max_fields >= 1");
    aeo_plan_s* p = new (std::nothrow) aeo_plan_s();
    AEO_REQUIRE(p != nullptr, AEO_ERR_HIP, "aeo_plan_create_points: out of host memory");
    p->kind = kPoints;
    p->L = l_max;
    p->B = n_basis(l_max);
    p->npts = npoints;
    p->max_fields = max_fields;
    p->nchunks = (npoints + kPtsChunk - 1) / kPtsChunk;
    std::vector<double> w(npoints);
    for (int i = 0; i < npoints; ++i) w[i] = weights ? weights[i] : 4.0 * M_PI / npoints;
    double *d_lat = nullptr, *d_lon = nullptr;
    const int nwork = std::max(max_fields, 4);
    int rc = AEO_OK;
    do {
        if ((rc = upload(&d_lat, lat, (size_t)npoints))) break;
        if ((rc = upload(&d_lon, lon, (size_t)npoints))) break;
        if ((rc = upload(&p->d_w, w.data(), w.size()))) break;
        if ((rc = upload(&p->d_pcoslat, coslat, (size_t)npoints))) break;
        if ((rc = dev_alloc(&p->d_Yk, (size_t)2 * p->B * npoints))) break;
        if ((rc = dev_alloc(&p->d_Yp, (size_t)p->B * npoints))) break;
        if ((rc = dev_alloc(&p->d_gw, (size_t)nwork * npoints))) break;
        if ((rc = dev_alloc(&p->d_part, (size_t)p->nchunks * nwork * p->B))) break;
        if ((rc = common_tables(p))) break;
        if ((rc = points_build_tables(p, d_lat, d_lon, nullptr))) break;
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            set_error(std::string("aeo_plan_create_points: ") + hipGetErrorString(e));
            rc = AEO_ERR_HIP;
        }
    } while (0);
    dev_free(d_lat);
    dev_free(d_lon);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return AEO_OK;
}

static int enable_split_points(aeo_plan_s* plan, int nslices) {
    const int B = plan->B;
    const int K2 = rup(2 * B, 256);
    AEO_REQUIRE(split_exact(nslices, K2) && split_exact(nslices, kPtsChunk), AEO_ERR_UNSUPPORTED,
                "aeo_plan_enable_split: INT32 accumulation would not be exact at this size");
    if (plan->nslices == nslices) return AEO_OK;
    ++plan->tables_gen;
    for (void* q : {(void*)plan->d_Ysl, (void*)plan->d_eYs, (void*)plan->d_Yasl,
                    (void*)plan->d_eYa})
        dev_free(q);
    plan->d_Ysl = nullptr; plan->d_eYs = nullptr; plan->d_Yasl = nullptr; plan->d_eYa = nullptr;
    plan->nslices = 0;
    plan->nprow = rup(plan->npts, 16);
    plan->K2 = K2;
    plan->R2p = rup(2 * B, 16);
    plan->Npad = plan->nchunks * kPtsChunk;
    int rc;
    if ((rc = dev_alloc(&plan->d_Ysl, (size_t)nslices * plan->nprow * plan->K2))) return rc;
    if ((rc = dev_alloc(&plan->d_eYs, (size_t)plan->nprow))) return rc;
    if ((rc = dev_alloc(&plan->d_Yasl, (size_t)nslices * plan->R2p * plan->Npad))) return rc;
    if ((rc = dev_alloc(&plan->d_eYa, (size_t)plan->R2p))) return rc;
    if ((rc = points_build_split_tables(plan, nslices, nullptr))) return rc;
    AEO_CHECK_HIP(hipDeviceSynchronize());
    plan->nslices = nslices;
    return AEO_OK;
}

int aeo_plan_enable_split(aeo_plan_t plan, int nslices) {
    AEO_REQUIRE(plan != nullptr, AEO_ERR_INVALID, "aeo_plan_enable_split: plan is NULL");
    AEO_REQUIRE(nslices >= 2 && nslices <= 8, AEO_ERR_INVALID,
                "aeo_plan_enable_split: nslices must be in [2, 8]");
    if (plan->kind == kPoints) return enable_split_points(plan, nslices);
    const int L = plan->L, nh = plan->nh;
    AEO_REQUIRE(split_exact(nslices, split_kp(L, 0)) && split_exact(nslices, nh),
                AEO_ERR_UNSUPPORTED,
                "aeo_plan_enable_split: INT32 accumulation would not be exact at this size");
    if (plan->nslices == nslices) return AEO_OK;
    ++plan->tables_gen;
    for (void* q : {(void*)plan->d_Psl, (void*)plan->d_offs_ss, (void*)plan->d_esyn,
                    (void*)plan->d_Pasl, (void*)plan->d_offs_sa, (void*)plan->d_eana,
                    (void*)plan->d_offs_ea})
        dev_free(q);
    plan->d_Psl = nullptr; plan->d_offs_ss = nullptr; plan->d_esyn = nullptr;
    plan->d_Pasl = nullptr; plan->d_offs_sa = nullptr; plan->d_eana = nullptr;
    plan->d_offs_ea = nullptr;
    plan->nslices = 0;
    const int nhp = rup(nh, 16), nhk = split_nhk(nh);
    std::vector<long long> oss(L + 2, 0), osa(L + 2, 0), oea(L + 2, 0);
    for (int m = 0; m <= L; ++m) {
        const int kp = split_kp(L, m);
        oss[m + 1] = oss[m] + 2LL * nslices * nhp * split_kpad(L, m);
        osa[m + 1] = osa[m] + 2LL * nslices * rup(kp, 16) * nhk;
        oea[m + 1] = oea[m] + 2LL * rup(kp, 16);
    }
    int rc;
    if ((rc = dev_alloc(&plan->d_Psl, (size_t)oss[L + 1]))) return rc;
    if ((rc = dev_alloc(&plan->d_Pasl, (size_t)osa[L + 1]))) return rc;
    if ((rc = dev_alloc(&plan->d_esyn, (size_t)(L + 1) * 2 * nhp))) return rc;
    if ((rc = dev_alloc(&plan->d_eana, (size_t)oea[L + 1]))) return rc;
    if ((rc = upload(&plan->d_offs_ss, oss.data(), oss.size()))) return rc;
    if ((rc = upload(&plan->d_offs_sa, osa.data(), osa.size()))) return rc;
    if ((rc = upload(&plan->d_offs_ea, oea.data(), oea.size()))) return rc;
    if ((rc = gauss_build_split_tables(plan, nslices, nullptr))) return rc;
    AEO_CHECK_HIP(hipDeviceSynchronize());
    plan->nslices = nslices;
    return AEO_OK;
}

int aeo_model_set_precision(aeo_model_t model, int nslices) {
    AEO_REQUIRE(model != nullptr, AEO_ERR_INVALID, "aeo_model_set_precision: model is NULL");
    AEO_REQUIRE(nslices == 0 || model->plan->nslices == nslices,
                AEO_ERR_STATE,
                "aeo_model_set_precision: enable the split tables on the plan first "
                "(aeo_plan_enable_split with the same nslices)");
    model->nslices = nslices;
    return AEO_OK;
}

int aeo_model_set_tiles(aeo_model_t model, int members_per_block) {
    AEO_REQUIRE(model != nullptr, AEO_ERR_INVALID, "aeo_model_set_tiles: model is NULL");
    AEO_REQUIRE(members_per_block == 0 || members_per_block == 1 || members_per_block == 2 ||
                    members_per_block == 4 || members_per_block == 8,
                AEO_ERR_INVALID,
                "aeo_model_set_tiles: members_per_block must be 0 (GEMM form), 1, 2, 4 or 8");
    if (members_per_block == 0) {
        int rc = gemm_stage_setup(model);
        if (rc) return rc;
    }
    model->mb = members_per_block;
    return AEO_OK;
}

int aeo_model_set_gemm_tiles(aeo_model_t model, int syn_cols, int epi_members, int syn_ntiles) {
    AEO_REQUIRE(model != nullptr, AEO_ERR_INVALID, "aeo_model_set_gemm_tiles: model is NULL");
    return gemm_stage_tiles(model, syn_cols, epi_members, syn_ntiles);
}

int aeo_plan_destroy(aeo_plan_t plan) {
    free_plan(plan);
    return AEO_OK;
}

int aeo_plan_info(aeo_plan_t plan, int* l_max, int* npoints, int* nbasis) {
    AEO_REQUIRE(plan != nullptr, AEO_ERR_INVALID, "aeo_plan_info: plan is NULL");
    if (l_max) *l_max = plan->L;
    if (npoints) *npoints = plan->kind == kGauss ? plan->nlat * plan->nlon : plan->npts;
    if (nbasis) *nbasis = plan->B;
    return AEO_OK;
}

int aeo_plan_basis(aeo_plan_t plan, double* Y_out, void* stream) {
    AEO_REQUIRE(plan && Y_out, AEO_ERR_INVALID, "aeo_plan_basis: NULL argument");
    AEO_REQUIRE(plan->kind == kPoints, AEO_ERR_UNSUPPORTED,
                "aeo_plan_basis: Gauss plans do not materialise the dense basis");
    AEO_CHECK_HIP(hipMemcpyAsync(Y_out, plan->d_Yp, sizeof(double2) * plan->B * plan->npts,
                                 hipMemcpyDeviceToDevice, as_stream(stream)));
    return AEO_OK;
}

int aeo_analysis(aeo_plan_t plan, const double* f, int nfields, double* a_out, void* stream) {
    AEO_REQUIRE(plan && f && a_out, AEO_ERR_INVALID, "aeo_analysis: NULL argument");
    AEO_REQUIRE(nfields >= 1, AEO_ERR_INVALID, "aeo_analysis: nfields must be >= 1");
    if (plan->kind == kGauss)
        return gauss_analysis(plan, f, nfields, reinterpret_cast<double2*>(a_out),
                              as_stream(stream));
    return points_analysis(plan, f, nfields, reinterpret_cast<double2*>(a_out), as_stream(stream));
}

int aeo_synthesis(aeo_plan_t plan, const double* a, int nfields, double* f_out, void* stream) {
    AEO_REQUIRE(plan && a && f_out, AEO_ERR_INVALID, "aeo_synthesis: NULL argument");
    AEO_REQUIRE(nfields >= 1, AEO_ERR_INVALID, "aeo_synthesis: nfields must be >= 1");
    if (plan->kind == kGauss)
        return gauss_synthesis(plan, reinterpret_cast<const double2*>(a), nfields, f_out,
                               as_stream(stream));
    return points_synthesis(plan, reinterpret_cast<const double2*>(a), nfields, f_out,
                            as_stream(stream));
}

int aeo_spec_op(int op, int l_max, double radius, const double* in, double* out, int nfields,
                void* stream) {
    AEO_REQUIRE(in && out, AEO_ERR_INVALID, "aeo_spec_op: NULL argument");
    AEO_REQUIRE(l_max >= 0 && nfields >= 1, AEO_ERR_INVALID, "aeo_spec_op: bad sizes");
    AEO_REQUIRE(op >= AEO_OP_LAP && op <= AEO_OP_INVLAP_ZERO, AEO_ERR_UNSUPPORTED,
                "aeo_spec_op: unknown operator code");
    AEO_REQUIRE(radius > 0.0 && std::isfinite(radius), AEO_ERR_INVALID,
                "aeo_spec_op: radius must be finite and > 0");
    AEO_REQUIRE(!((op == AEO_OP_SINTH_DTH || op == AEO_OP_SINTH_DTH_T) && in == out),
                AEO_ERR_INVALID, "aeo_spec_op: the l+-1 coupling cannot run in place");
    return launch_spec_op(op, l_max, radius, reinterpret_cast<const double2*>(in),
                          reinterpret_cast<double2*>(out), nfields, as_stream(stream));
}

int aeo_jacobian(aeo_plan_t prod, double radius, const double* a, const double* b, double* out,
                 int dealias, void* stream) {
    AEO_REQUIRE(prod && a && b && out, AEO_ERR_INVALID, "aeo_jacobian: NULL argument");
    AEO_REQUIRE(radius > 0.0, AEO_ERR_INVALID, "aeo_jacobian: radius must be > 0");
    auto A = reinterpret_cast<const double2*>(a);
    auto Bp = reinterpret_cast<const double2*>(b);
    auto O = reinterpret_cast<double2*>(out);
    if (prod->kind == kGauss) return gauss_jacobian(prod, radius, A, Bp, O, dealias, as_stream(stream));
    return points_jacobian(prod, radius, A, Bp, O, dealias, as_stream(stream));
}

static int velocity_common(aeo_plan_t plan, double radius, const double* q, double* o1,
                           double* o2, int mode, void* stream) {
    AEO_REQUIRE(plan && q && o1 && o2, AEO_ERR_INVALID, "velocity/gradient: NULL argument");
    AEO_REQUIRE(radius > 0.0, AEO_ERR_INVALID, "velocity/gradient: radius must be > 0");
    hipStream_t s = as_stream(stream);
    const int L = plan->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    // spec[0] = d/dlambda q, spec[1] = S q; the synthesised fields land in
    // o1 (S q) and o2 (q_lambda) before the pointwise stage rewrites them.
    double2* spec = plan->d_spec;
    int rc = launch_spec_op(AEO_OP_DLAMBDA, L, radius, reinterpret_cast<const double2*>(q), spec,
                            1, s);
    if (rc) return rc;
    rc = launch_spec_op(AEO_OP_SINTH_DTH, L, radius, reinterpret_cast<const double2*>(q),
                        spec + fs, 1, s);
    if (rc) return rc;
    const int npts = plan->kind == kGauss ? plan->nlat * plan->nlon : plan->npts;
    rc = aeo_synthesis(plan, reinterpret_cast<const double*>(spec + fs), 1, o1, stream);
    if (rc) return rc;
    rc = aeo_synthesis(plan, reinterpret_cast<const double*>(spec), 1, o2, stream);
    if (rc) return rc;
    const double* c = plan->kind == kGauss ? plan->d_coslat : plan->d_pcoslat;
    const int row_len = plan->kind == kGauss ? plan->nlon : 0;
    return launch_velocity_points(npts, row_len, radius, c, o1, o2, o1, o2, mode, s);
}

int aeo_velocity(aeo_plan_t plan, double radius, const double* psi, double* u, double* v,
                 void* stream) {
    return velocity_common(plan, radius, psi, u, v, 0, stream);
}

int aeo_gradient(aeo_plan_t plan, double radius, const double* q, double* dqdx, double* dqdy,
                 void* stream) {
    // velocity_common writes (S q, q_lambda) into (o1, o2) and mode 1 maps
    // them to (q_lambda / c, -S q / (R c)).
    return velocity_common(plan, radius, q, dqdx, dqdy, 1, stream);
}

static int model_common(aeo_plan_t prod, double radius, aeo_model_s* m) {
    const int L = prod->L;
    m->plan = prod;
    m->L = L;
    m->R = radius;
    std::vector<double> lap(L + 1), nulap(L + 1), nu4lap2(L + 1);
    for (int l = 0; l <= L; ++l) {
        lap[l] = -(double)l * (l + 1.0) / (radius * radius);
        const double le = (l == 0) ? 1.0 / (radius * radius) : lap[l];
        nulap[l] = m->nu * le;
        nu4lap2[l] = m->nu4 * (lap[l] * lap[l]);
    }
    int rc;
    if ((rc = upload(&m->d_lap, lap.data(), lap.size()))) return rc;
    if ((rc = upload(&m->d_nulap, nulap.data(), nulap.size()))) return rc;
    if ((rc = upload(&m->d_nu4lap2, nu4lap2.data(), nu4lap2.size()))) return rc;
    const int nsyn = (m->kind == kBVE) ? 4 : 12;
    const int nana = (m->kind == kBVE) ? 1 : 4;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    const size_t nm = (size_t)m->nmem;
    if (prod->kind == kGauss) {
        const size_t plane = (size_t)prod->nh * (L + 1);
        if ((rc = dev_alloc(&m->d_E, plane * nsyn * nm))) return rc;
        if ((rc = dev_alloc(&m->d_O, plane * nsyn * nm))) return rc;
        if (m->kind == kBVE && nm == 1) {
            if ((rc = dev_alloc(&m->d_E1, plane * nsyn))) return rc;
            if ((rc = dev_alloc(&m->d_O1, plane * nsyn))) return rc;
            if ((rc = bve1_prepare(prod, nullptr))) return rc;
        }
        if ((rc = dev_alloc(&m->d_S, plane * nana * nm))) return rc;
        if ((rc = dev_alloc(&m->d_D, plane * nana * nm))) return rc;
    } else {
        if ((rc = dev_alloc(&m->d_gw, (size_t)prod->npts * nana))) return rc;
        if ((rc = dev_alloc(&m->d_part, (size_t)prod->nchunks * nana * prod->B))) return rc;
    }
    const size_t st = fs * m->nstate * nm;
    if ((rc = dev_alloc(&m->d_ystage, 2 * st))) return rc;
    if ((rc = dev_alloc(&m->d_ybuf, 2 * st))) return rc;
    if ((rc = dev_alloc(&m->d_ksum, st))) return rc;
    if (prod->kind == kGauss && m->kind == kBVE) {
        if ((rc = dev_alloc(&m->d_fint, st))) return rc;
    }
    return AEO_OK;
}

static int default_mb(int nmem) {
    int mb = 1;
    while (mb * 2 <= nmem && mb < 4) mb *= 2;
    return mb;
}

int aeo_bve_create_batch(aeo_plan_t prod, double radius, double omega, double viscosity,
                         int nmembers, aeo_model_t* out) {
    AEO_REQUIRE(out != nullptr, AEO_ERR_INVALID, "aeo_bve_create: out is NULL");
    *out = nullptr;
    AEO_REQUIRE(prod != nullptr, AEO_ERR_INVALID, "aeo_bve_create: plan is NULL");
    AEO_REQUIRE(radius > 0.0 && std::isfinite(radius), AEO_ERR_INVALID,
                "aeo_bve_create: radius must be finite and > 0");
    AEO_REQUIRE(std::isfinite(omega) && std::isfinite(viscosity), AEO_ERR_INVALID,
                "aeo_bve_create: omega and viscosity must be finite");
    AEO_REQUIRE(nmembers >= 1, AEO_ERR_INVALID, "aeo_bve_create: nmembers must be >= 1");
    AEO_REQUIRE(nmembers == 1 || prod->kind == kGauss, AEO_ERR_UNSUPPORTED,
                "aeo_bve_create: member batches need a Gauss plan");
    aeo_model_s* m = new (std::nothrow) aeo_model_s();
    AEO_REQUIRE(m != nullptr, AEO_ERR_HIP, "aeo_bve_create: out of host memory");
    m->kind = kBVE;
    m->nstate = 1;
    m->omega = omega;
    m->nu = viscosity;
    m->nmem = nmembers;
    m->mb = default_mb(nmembers);
    int rc = model_common(prod, radius, m);
    if (rc) {
        free_model(m);
        return rc;
    }
    *out = m;
    return AEO_OK;
}

int aeo_bve_create(aeo_plan_t prod, double radius, double omega, double viscosity,
                   aeo_model_t* out) {
    return aeo_bve_create_batch(prod, radius, omega, viscosity, 1, out);
}

int aeo_swe_create(aeo_plan_t prod, double radius, double omega, double gravity,
                   double mean_depth, double hyperdiffusion_nu4, const double* phi_s_lm,
                   aeo_model_t* out) {
    AEO_REQUIRE(out != nullptr, AEO_ERR_INVALID, "aeo_swe_create: out is NULL");
    *out = nullptr;
    AEO_REQUIRE(prod != nullptr, AEO_ERR_INVALID, "aeo_swe_create: plan is NULL");
    AEO_REQUIRE(std::isfinite(gravity) && gravity > 0, AEO_ERR_INVALID,
                "gravity must be finite and > 0");
    AEO_REQUIRE(std::isfinite(mean_depth) && mean_depth > 0, AEO_ERR_INVALID,
                "mean_depth must be finite and > 0");
    AEO_REQUIRE(std::isfinite(hyperdiffusion_nu4) && hyperdiffusion_nu4 >= 0, AEO_ERR_INVALID,
                "hyperdiffusion_nu4 must be finite and >= 0");
    AEO_REQUIRE(radius > 0.0 && std::isfinite(radius) && std::isfinite(omega), AEO_ERR_INVALID,
                "aeo_swe_create: radius must be finite and > 0, omega finite");
    aeo_model_s* m = new (std::nothrow) aeo_model_s();
    AEO_REQUIRE(m != nullptr, AEO_ERR_HIP, "aeo_swe_create: out of host memory");
    m->kind = kSWE;
    m->nstate = 3;
    m->nmem = 1;
    m->mb = 1;
    m->omega = omega;
    m->gravity = gravity;
    m->mean_depth = mean_depth;
    m->phi0 = gravity * mean_depth;
    m->nu4 = hyperdiffusion_nu4;
    int rc = model_common(prod, radius, m);
    if (!rc && phi_s_lm) {
        // lap * phi_s and nu4 * lap^2 * (phi_s with the (0,0) entry zeroed)
        // (phy/shallow_water.py:240-253), formed once on the host.
        const int L = prod->L;
        const size_t fs = (size_t)(L + 1) * (L + 1);
        std::vector<double2> ps(fs), lps(fs), nps(fs);
        if (hipMemcpy(ps.data(), phi_s_lm, fs * sizeof(double2), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            set_error("aeo_swe_create: cannot read phi_s_lm");
            rc = AEO_ERR_HIP;
        } else {
            bool any = false;
            for (size_t i = 0; i < fs; ++i) any = any || ps[i].x != 0.0 || ps[i].y != 0.0;
            m->has_topo = any;
            for (int l = 0; l <= L; ++l) {
                const double lap = -(double)l * (l + 1.0) / (radius * radius);
                for (int mm = 0; mm <= L; ++mm) {
                    const size_t i = (size_t)l * (L + 1) + mm;
                    lps[i] = make_double2(lap * ps[i].x, lap * ps[i].y);
                    const double2 an = (i == 0) ? make_double2(0.0, 0.0) : ps[i];
                    const double q = hyperdiffusion_nu4 * (lap * lap);
                    nps[i] = make_double2(q * an.x, q * an.y);
                }
            }
            if (any) {
                rc = upload(&m->d_lap_phi_s, lps.data(), fs);
                if (!rc) rc = upload(&m->d_nu4_lap2_phi_s, nps.data(), fs);
            }
        }
    }
    if (rc) {
        free_model(m);
        return rc;
    }
    *out = m;
    return AEO_OK;
}

int aeo_model_destroy(aeo_model_t model) {
    free_model(model);
    return AEO_OK;
}

// The Gauss model kernels keep states in the internal [member][field][m][l]
// layout (columns of one m contiguous); the boundary layout is [..][l][m].
static int to_internal(aeo_model_s* m, const double2* user, double2* internal, hipStream_t s) {
    return launch_transpose(user, internal, m->L + 1, m->nstate * m->nmem, s);
}

int aeo_model_tendency(aeo_model_t model, const double* y, const double* forcing, double* out,
                       void* stream) {
    AEO_REQUIRE(model && y && out, AEO_ERR_INVALID, "aeo_model_tendency: NULL argument");
    AEO_REQUIRE(!(forcing && model->kind == kSWE), AEO_ERR_UNSUPPORTED,
                "aeo_model_tendency: forcing is only defined for the BVE model");
    hipStream_t s = as_stream(stream);
    auto Y = reinterpret_cast<const double2*>(y);
    auto F = reinterpret_cast<const double2*>(forcing);
    auto K = reinterpret_cast<double2*>(out);
    if (model->plan->kind == kGauss) {
        const size_t st = (size_t)(model->L + 1) * (model->L + 1) * model->nstate * model->nmem;
        double2* yi = model->d_ystage;
        double2* ki = model->d_ystage + st;
        int rc = to_internal(model, Y, yi, s);
        if (!rc && F) rc = to_internal(model, F, model->d_fint, s);
        if (!rc) rc = gauss_model_prime(model, yi, s);
        if (!rc) rc = gauss_model_stage(model, yi, yi, F ? model->d_fint : nullptr, ki, -1, 0.0,
                                        nullptr, false, s);
        if (!rc) rc = launch_transpose(ki, K, model->L + 1, model->nstate * model->nmem, s);
        return rc;
    }
    return points_model_stage(model, Y, Y, F, K, -1, 0.0, nullptr, s);
}

int aeo_model_rk4(aeo_model_t model, const double* y, double dt, int nsteps,
                  const double* forcing, double* y_out, void* stream) {
    AEO_REQUIRE(model && y && y_out, AEO_ERR_INVALID, "aeo_model_rk4: NULL argument");
    AEO_REQUIRE(nsteps >= 0, AEO_ERR_INVALID, "aeo_model_rk4: nsteps must be >= 0");
    AEO_REQUIRE(std::isfinite(dt), AEO_ERR_INVALID, "aeo_model_rk4: dt must be finite");
    AEO_REQUIRE(!(forcing && model->kind == kSWE), AEO_ERR_UNSUPPORTED,
                "aeo_model_rk4: forcing is only defined for the BVE model");
    hipStream_t s = as_stream(stream);
    const size_t st = (size_t)(model->L + 1) * (model->L + 1) * model->nstate * model->nmem;
    auto F = reinterpret_cast<const double2*>(forcing);
    if (nsteps == 0) {
        if (y != y_out)
            AEO_CHECK_HIP(hipMemcpyAsync(y_out, y, st * sizeof(double2), hipMemcpyDeviceToDevice, s));
        return AEO_OK;
    }
    const bool gauss = model->plan->kind == kGauss;
    double2* ystage[2] = {model->d_ystage, model->d_ystage + st};
    double2* ybuf[2] = {model->d_ybuf, model->d_ybuf + st};
    const double2* cur = reinterpret_cast<const double2*>(y);
    if (gauss) {
        // internal copy of the input in ybuf[1]: step 0 reads it and writes ybuf[0]
        int rc = to_internal(model, cur, ybuf[1], s);
        if (!rc && F) rc = to_internal(model, F, model->d_fint, s);
        if (rc) return rc;
        cur = ybuf[1];
        F = F ? model->d_fint : nullptr;
        rc = gauss_model_prime(model, cur, s);
        if (rc) return rc;
    }
    for (int step = 0; step < nsteps; ++step) {
        double2* out = ybuf[step % 2];
        for (int k = 0; k < 4; ++k) {
            const double2* ys = (k == 0) ? cur : ystage[(k - 1) % 2];
            double2* ynext = (k == 3) ? out : ystage[k % 2];
            int rc;
            if (gauss) {
                const bool syn_next = !(step == nsteps - 1 && k == 3);
                rc = gauss_model_stage(model, ys, cur, F, nullptr, k, dt, ynext, syn_next, s);
            } else {
                rc = points_model_stage(model, ys, cur, F, nullptr, k, dt, ynext, s);
            }
            if (rc) return rc;
        }
        cur = out;
    }
    if (gauss) return launch_transpose(cur, reinterpret_cast<double2*>(y_out), model->L + 1,
                                       model->nstate * model->nmem, s);
    AEO_CHECK_HIP(hipMemcpyAsync(y_out, cur, st * sizeof(double2), hipMemcpyDeviceToDevice, s));
    return AEO_OK;
}

// ---------------------------------------------------------------------------
// Checkpointed forward run and the reverse (adjoint) sweep of the BVE step.
// ---------------------------------------------------------------------------
static size_t state_elems(const aeo_model_s* m) {
    return (size_t)(m->L + 1) * (m->L + 1) * m->nstate * m->nmem;
}

static int adjoint_supported(const aeo_model_s* m, const char* who) {
    if (m->kind != kBVE || m->plan->kind != kGauss) {
        set_error(std::string(who) + ": the adjoint is implemented for BVE models on Gauss plans");
        return AEO_ERR_UNSUPPORTED;
    }
    return AEO_OK;
}

static bool adjoint_ready(const aeo_model_s* m) {
    return m->d_aE && m->d_aO && m->d_aS && m->d_aD && m->d_asum && m->d_abuf && m->d_kbuf;
}

static void free_adjoint(aeo_model_s* m) {
    for (double2** q : {&m->d_aE, &m->d_aO, &m->d_aS, &m->d_aD, &m->d_asum, &m->d_abuf,
                        &m->d_kbuf}) {
        dev_free(*q);
        *q = nullptr;
    }
}

// The adjoint execute calls allocate nothing: they need the workspaces this
// set-up call allocates (all or none).
static int require_adjoint(const aeo_model_s* m, const char* who) {
    if (adjoint_ready(m)) return AEO_OK;
    set_error(std::string(who) + ": call aeo_model_enable_adjoint on the model first");
    return AEO_ERR_STATE;
}

int aeo_model_enable_adjoint(aeo_model_t model) {
    AEO_REQUIRE(model != nullptr, AEO_ERR_INVALID, "aeo_model_enable_adjoint: model is NULL");
    int rc = adjoint_supported(model, "aeo_model_enable_adjoint");
    if (rc) return rc;
    if (adjoint_ready(model)) return AEO_OK;
    free_adjoint(model);
    const size_t plane = (size_t)model->plan->nh * (model->L + 1) * model->nmem;
    const size_t st = state_elems(model);
    if (!(rc = dev_alloc(&model->d_aE, plane * 5)) && !(rc = dev_alloc(&model->d_aO, plane * 5)) &&
        !(rc = dev_alloc(&model->d_aS, plane * 4)) && !(rc = dev_alloc(&model->d_aD, plane * 4)) &&
        !(rc = dev_alloc(&model->d_asum, st)) && !(rc = dev_alloc(&model->d_abuf, 2 * st)))
        rc = dev_alloc(&model->d_kbuf, 2 * st);
    if (rc) free_adjoint(model);
    return rc;
}

int aeo_model_checkpoint_elems(aeo_model_t model, int nsteps, long long* out) {
    AEO_REQUIRE(model && out, AEO_ERR_INVALID, "aeo_model_checkpoint_elems: NULL argument");
    AEO_REQUIRE(nsteps >= 0, AEO_ERR_INVALID, "aeo_model_checkpoint_elems: nsteps must be >= 0");
    *out = (long long)nsteps * 4 * (long long)state_elems(model);
    return AEO_OK;
}

int aeo_model_rk4_record(aeo_model_t model, const double* y, double dt, int nsteps,
                         const double* forcing, double* y_out, double* checkpoints,
                         void* stream) {
    AEO_REQUIRE(model && y && y_out, AEO_ERR_INVALID, "aeo_model_rk4_record: NULL argument");
    AEO_REQUIRE(nsteps >= 0 && std::isfinite(dt), AEO_ERR_INVALID,
                "aeo_model_rk4_record: need nsteps >= 0 and a finite dt");
    AEO_REQUIRE(nsteps == 0 || checkpoints, AEO_ERR_INVALID,
                "aeo_model_rk4_record: checkpoints is NULL");
    int rc = adjoint_supported(model, "aeo_model_rk4_record");
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const size_t st = state_elems(model);
    if (nsteps == 0) {
        if (y != y_out)
            AEO_CHECK_HIP(hipMemcpyAsync(y_out, y, st * sizeof(double2), hipMemcpyDeviceToDevice, s));
        return AEO_OK;
    }
    // the stage inputs are written straight into their checkpoint slots:
    // slot (step, k) = input state of RK4 stage k of that step
    double2* ck = reinterpret_cast<double2*>(checkpoints);
    auto slot = [&](int step, int k) { return ck + ((size_t)step * 4 + k) * st; };
    const double2* F = reinterpret_cast<const double2*>(forcing);
    rc = to_internal(model, reinterpret_cast<const double2*>(y), slot(0, 0), s);
    if (!rc && F) rc = to_internal(model, F, model->d_fint, s);
    if (!rc) rc = gauss_model_prime(model, slot(0, 0), s);
    if (rc) return rc;
    F = F ? model->d_fint : nullptr;
    double2* last = model->d_ybuf;
    for (int step = 0; step < nsteps; ++step) {
        for (int k = 0; k < 4; ++k) {
            double2* ynext = (k < 3) ? slot(step, k + 1) : (step + 1 < nsteps ? slot(step + 1, 0) : last);
            const bool syn_next = !(step == nsteps - 1 && k == 3);
            rc = gauss_model_stage(model, slot(step, k), slot(step, 0), F, nullptr, k, dt, ynext,
                                   syn_next, s);
            if (rc) return rc;
        }
    }
    return launch_transpose(last, reinterpret_cast<double2*>(y_out), model->L + 1,
                            model->nstate * model->nmem, s);
}

int aeo_model_rk4_adjoint(aeo_model_t model, const double* checkpoints, double dt, int nsteps,
                          const double* ybar_out, double* ybar, void* stream) {
    AEO_REQUIRE(model && ybar_out && ybar, AEO_ERR_INVALID, "aeo_model_rk4_adjoint: NULL argument");
    AEO_REQUIRE(nsteps >= 0 && std::isfinite(dt), AEO_ERR_INVALID,
                "aeo_model_rk4_adjoint: need nsteps >= 0 and a finite dt");
    AEO_REQUIRE(nsteps == 0 || checkpoints, AEO_ERR_INVALID,
                "aeo_model_rk4_adjoint: checkpoints is NULL");
    int rc = adjoint_supported(model, "aeo_model_rk4_adjoint");
    if (!rc) rc = require_adjoint(model, "aeo_model_rk4_adjoint");
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const size_t st = state_elems(model);
    if (nsteps == 0) {
        if (ybar_out != ybar)
            AEO_CHECK_HIP(hipMemcpyAsync(ybar, ybar_out, st * sizeof(double2),
                                         hipMemcpyDeviceToDevice, s));
        return AEO_OK;
    }
    const double2* ck = reinterpret_cast<const double2*>(checkpoints);
    auto slot = [&](int step, int k) { return ck + ((size_t)step * 4 + k) * st; };
    double2* abuf[2] = {model->d_abuf, model->d_abuf + st};
    double2* kbuf[2] = {model->d_kbuf, model->d_kbuf + st};
    rc = to_internal(model, reinterpret_cast<const double2*>(ybar_out), abuf[0], s);
    if (!rc) rc = launch_scale(abuf[0], kbuf[0], dt / 6.0, st, s);
    if (!rc) rc = gauss_adjoint_prime(model, slot(nsteps - 1, 3), kbuf[0], s);
    if (rc) return rc;
    int ai = 0, ki = 0;
    for (int step = nsteps - 1; step >= 0; --step) {
        for (int r = 0; r < 4; ++r) {
            const double2* fwd_next =
                (r < 3) ? slot(step, 2 - r) : (step > 0 ? slot(step - 1, 3) : nullptr);
            rc = gauss_adjoint_stage(model, kbuf[ki], r, dt, abuf[ai], abuf[ai ^ 1], nullptr,
                                     kbuf[ki ^ 1], fwd_next, s);
            if (rc) return rc;
            ki ^= 1;
        }
        ai ^= 1;
    }
    return launch_transpose(abuf[ai], reinterpret_cast<double2*>(ybar), model->L + 1,
                            model->nstate * model->nmem, s);
}

int aeo_model_tendency_adjoint(aeo_model_t model, const double* y, const double* kbar,
                               double* ybar, void* stream) {
    AEO_REQUIRE(model && y && kbar && ybar, AEO_ERR_INVALID,
                "aeo_model_tendency_adjoint: NULL argument");
    int rc = adjoint_supported(model, "aeo_model_tendency_adjoint");
    if (!rc) rc = require_adjoint(model, "aeo_model_tendency_adjoint");
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const size_t st = state_elems(model);
    double2* yi = model->d_abuf;
    double2* out = model->d_abuf + st;
    rc = to_internal(model, reinterpret_cast<const double2*>(y), yi, s);
    if (!rc) rc = to_internal(model, reinterpret_cast<const double2*>(kbar), model->d_kbuf, s);
    if (!rc) rc = gauss_adjoint_prime(model, yi, model->d_kbuf, s);
    if (!rc) rc = gauss_adjoint_stage(model, model->d_kbuf, -1, 0.0, nullptr, nullptr, out,
                                      nullptr, nullptr, s);
    if (!rc) rc = launch_transpose(out, reinterpret_cast<double2*>(ybar), model->L + 1,
                                   model->nstate * model->nmem, s);
    return rc;
}

int aeo_model_profile(aeo_model_t model, int kernel_class, int max_launches) {
    AEO_REQUIRE(model != nullptr, AEO_ERR_INVALID, "aeo_model_profile: model is NULL");
    AEO_REQUIRE(kernel_class >= -1 && kernel_class <= kClsAnalysis, AEO_ERR_INVALID,
                "aeo_model_profile: kernel_class must be -1 (off), 0, 1 or 2");
    free_prof(model->prof);
    if (kernel_class < 0 || max_launches <= 0) return AEO_OK;
    ProfRec& pr = model->prof;
    pr.ev = new (std::nothrow) hipEvent_t[2 * (size_t)max_launches];
    AEO_REQUIRE(pr.ev != nullptr, AEO_ERR_HIP, "aeo_model_profile: out of host memory");
    for (int i = 0; i < 2 * max_launches; ++i) {
        hipError_t e = hipEventCreate(&pr.ev[i]);
        if (e != hipSuccess) {
            pr.cap = i / 2;
            for (int j = 2 * pr.cap; j < i; ++j) (void)hipEventDestroy(pr.ev[j]);
            free_prof(pr);
            set_error(std::string("aeo_model_profile: hipEventCreate: ") + hipGetErrorString(e));
            return AEO_ERR_HIP;
        }
    }
    pr.cap = max_launches;
    pr.cls = kernel_class;
    pr.n = 0;
    return AEO_OK;
}

int aeo_model_profile_read(aeo_model_t model, double* total_ms, int* launches) {
    AEO_REQUIRE(model && total_ms && launches, AEO_ERR_INVALID,
                "aeo_model_profile_read: NULL argument");
    ProfRec& pr = model->prof;
    double tot = 0.0;
    for (int i = 0; i < pr.n; ++i) {
        AEO_CHECK_HIP(hipEventSynchronize(pr.ev[2 * i + 1]));
        float ms = 0.0f;
        AEO_CHECK_HIP(hipEventElapsedTime(&ms, pr.ev[2 * i], pr.ev[2 * i + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = pr.n;
    pr.n = 0;
    return AEO_OK;
}

int aeo_spectral_diag(int l_max, double radius, double omega, const double* zeta, int nfields,
                      double* out, void* stream) {
    AEO_REQUIRE(zeta && out, AEO_ERR_INVALID, "aeo_spectral_diag: NULL argument");
    AEO_REQUIRE(l_max >= 0 && nfields >= 1, AEO_ERR_INVALID, "aeo_spectral_diag: bad sizes");
    return launch_spectral_diag(l_max, radius, omega, reinterpret_cast<const double2*>(zeta),
                                nfields, out, as_stream(stream));
}

int aeo_grid_stats(int mode, int npts, const double* w, const double* a, const double* b,
                   const double* c, const double* d, const double* k, double* out, void* stream) {
    AEO_REQUIRE(out && a && npts >= 1, AEO_ERR_INVALID, "aeo_grid_stats: bad arguments");
    AEO_REQUIRE(mode == AEO_GRID_STATS_BVE || mode == AEO_GRID_STATS_SWE ||
                    mode == AEO_GRID_STATS_MINMAX,
                AEO_ERR_INVALID, "aeo_grid_stats: unknown mode");
    if (mode == AEO_GRID_STATS_BVE)
        AEO_REQUIRE(w && b && c, AEO_ERR_INVALID, "aeo_grid_stats: BVE needs w, zeta, u, v");
    if (mode == AEO_GRID_STATS_SWE) {
        AEO_REQUIRE(w && b && c && k, AEO_ERR_INVALID,
                    "aeo_grid_stats: SWE needs w, u, v, phi and the constants k");
        AEO_REQUIRE(k[2] != 0.0, AEO_ERR_INVALID, "aeo_grid_stats: gravity must be nonzero");
    }
    const double k0 = k ? k[0] : 0.0, k1 = k ? k[1] : 0.0, k2 = k ? k[2] : 1.0;
    return launch_grid_stats(mode, npts, w, a, b, c, d, k0, k1, k2, out, as_stream(stream));
}

}  // extern "C"
