// State files (host code): io_load_state / io_save_state
// (program/io.F90:349-483, 1502-1621) in the netCDF classic format the
// reference writes (nf90_create with nf90_clobber: CDF-1; CDF-2, the 64-bit
// offset variant, is read as well).  The reference reaches the format through
// the netCDF library; this is a self-contained reader and writer of the parts
// a state file uses: dimensions, global and variable attributes, and
// non-record double / int variables, all big-endian.
//
// A coll variable is stored Fortran (N, H, 2), i.e. dims (ReIm, H, r) in the
// file's C order, with its harmonics in the m-major order of io_state_nh
// (program/io.F90:1256-1269): nh = L m - Mp (m-1) m / 2 + l - m_ with
// m = m_ / Mp, and attributes L, M, Mp of the run that wrote it.  Loading
// maps every (l, m_) of this run to that index (modes the file lacks stay
// zero) and, when the file's radial grid differs from this run's, takes
// each value by 7-point Lagrange interpolation with the weights of
// io_interp_wts / io_interp (program/io.F90:1371-1431).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "lsd_internal.h"

namespace lsd {

namespace {

constexpr int NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6;
constexpr unsigned NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C;

struct NcAtt {
    int type = 0;
    std::vector<double> num;  // int and double values
    std::string text;         // char values
};

struct NcVar {
    std::vector<int> dims;
    std::map<std::string, NcAtt> atts;
    int type = 0;
    unsigned long long begin = 0;
};

struct NcFile {
    int version = 0;
    std::vector<std::pair<std::string, unsigned long long>> dims;
    std::map<std::string, NcAtt> gatts;
    std::map<std::string, NcVar> vars;
    std::vector<unsigned char> bytes;
};

struct Reader {
    const std::vector<unsigned char>& b;
    size_t p = 0;
    bool bad = false;
    explicit Reader(const std::vector<unsigned char>& bytes) : b(bytes) {}
    unsigned long long u(int n) {
        if (p + n > b.size()) {
            bad = true;
            return 0;
        }
        unsigned long long v = 0;
        for (int k = 0; k < n; k++) v = (v << 8) | b[p + k];
        p += n;
        return v;
    }
    std::string name() {
        const size_t n = (size_t)u(4);
        if (bad || p + n > b.size()) {
            bad = true;
            return "";
        }
        std::string s(reinterpret_cast<const char*>(&b[p]), n);
        p += (n + 3) & ~(size_t)3;
        return s;
    }
};

double be_double(const unsigned char* q) {
    unsigned long long v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | q[k];
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

int type_size(int t) { return t == NC_DOUBLE ? 8 : t == NC_INT ? 4 : t == 3 ? 2 : 1; }

bool read_atts(Reader& r, std::map<std::string, NcAtt>& out) {
    const unsigned tag = (unsigned)r.u(4);
    const unsigned n = (unsigned)r.u(4);
    if (tag == 0 && n == 0) return !r.bad;
    if (tag != NC_ATTRIBUTE) return false;
    for (unsigned i = 0; i < n && !r.bad; i++) {
        std::string nm = r.name();
        NcAtt a;
        a.type = (int)r.u(4);
        const size_t cnt = (size_t)r.u(4);
        const size_t sz = (size_t)type_size(a.type) * cnt;
        if (r.p + sz > r.b.size()) return false;
        const unsigned char* q = &r.b[r.p];
        if (a.type == NC_CHAR) {
            a.text.assign(reinterpret_cast<const char*>(q), cnt);
        } else if (a.type == NC_DOUBLE) {
            for (size_t k = 0; k < cnt; k++) a.num.push_back(be_double(q + 8 * k));
        } else if (a.type == NC_INT) {
            for (size_t k = 0; k < cnt; k++) {
                const unsigned v = ((unsigned)q[4 * k] << 24) | ((unsigned)q[4 * k + 1] << 16) |
                                   ((unsigned)q[4 * k + 2] << 8) | q[4 * k + 3];
                a.num.push_back((double)(int)v);
            }
        }
        r.p += (sz + 3) & ~(size_t)3;
        out[nm] = a;
    }
    return !r.bad;
}

int parse(const std::string& path, NcFile& f) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        set_error("state file not found!");
        return LSD_E_ARG;
    }
    std::fseek(fp, 0, SEEK_END);
    const long len = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f.bytes.resize(len > 0 ? (size_t)len : 0);
    const size_t got = f.bytes.empty() ? 0 : std::fread(f.bytes.data(), 1, f.bytes.size(), fp);
    std::fclose(fp);
    if (got != f.bytes.size() || f.bytes.size() < 32 || f.bytes[0] != 'C' || f.bytes[1] != 'D' ||
        f.bytes[2] != 'F' || (f.bytes[3] != 1 && f.bytes[3] != 2)) {
        set_error("state file: not a netCDF classic (CDF-1 / CDF-2) file");
        return LSD_E_ARG;
    }
    f.version = f.bytes[3];
    Reader r(f.bytes);
    r.p = 4;
    r.u(4);  // numrecs: state files have no record variables
    unsigned tag = (unsigned)r.u(4), n = (unsigned)r.u(4);
    if (tag == NC_DIMENSION) {
        for (unsigned i = 0; i < n && !r.bad; i++) {
            std::string nm = r.name();
            f.dims.push_back({nm, r.u(4)});
        }
    } else if (tag != 0 || n != 0) {
        r.bad = true;
    }
    if (r.bad || !read_atts(r, f.gatts)) {
        set_error("state file: malformed header");
        return LSD_E_ARG;
    }
    tag = (unsigned)r.u(4);
    n = (unsigned)r.u(4);
    if (tag == NC_VARIABLE) {
        for (unsigned i = 0; i < n && !r.bad; i++) {
            std::string nm = r.name();
            NcVar v;
            const unsigned nd = (unsigned)r.u(4);
            for (unsigned k = 0; k < nd; k++) v.dims.push_back((int)r.u(4));
            if (!read_atts(r, v.atts)) r.bad = true;
            v.type = (int)r.u(4);
            r.u(4);  // vsize
            v.begin = r.u(f.version == 2 ? 8 : 4);
            f.vars[nm] = v;
        }
    } else if (tag != 0 || n != 0) {
        r.bad = true;
    }
    if (r.bad) {
        set_error("state file: malformed header");
        return LSD_E_ARG;
    }
    return LSD_OK;
}

// Shape and doubles of a double variable (C order).
bool get_doubles(const NcFile& f, const std::string& nm, std::vector<size_t>& shape, std::vector<double>& out) {
    auto it = f.vars.find(nm);
    if (it == f.vars.end() || it->second.type != NC_DOUBLE) return false;
    shape.clear();
    size_t cnt = 1;
    for (int d : it->second.dims) {
        if (d < 0 || d >= (int)f.dims.size()) return false;
        shape.push_back((size_t)f.dims[d].second);
        cnt *= (size_t)f.dims[d].second;
    }
    if (it->second.begin + 8 * cnt > f.bytes.size()) return false;
    out.resize(cnt);
    for (size_t k = 0; k < cnt; k++) out[k] = be_double(&f.bytes[it->second.begin + 8 * k]);
    return true;
}

int att_int(const NcVar& v, const char* nm, int dflt) {
    auto it = v.atts.find(nm);
    return (it == v.atts.end() || it->second.num.empty()) ? dflt : (int)it->second.num[0];
}

// io_state_nh: index of (l, m_) in a file written with (iL, iM, iMp), -1 if absent.
long state_nh(int l, int m_, int iL, int iM, int iMp) {
    if (iMp <= 0 || m_ % iMp != 0 || l >= iL || m_ >= iM * iMp) return -1;
    const long m = m_ / iMp;
    return (long)iL * m - (long)iMp * (m - 1) * m / 2 + l - m_;
}

// First row of the inverse of the n x n Taylor matrix A (column-major, lda
// 7), the interpolation weights io_interp takes from mes_mat_invert: the
// solution of A^T w = e_1, by Gaussian elimination with partial pivoting.
bool first_row_of_inverse(int n, const double* A, double* w) {
    double T[7][8];
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) T[i][j] = A[i * 7 + j];  // (A^T)(i, j) = A(j, i)
        T[i][n] = i == 0 ? 1.0 : 0.0;
    }
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int i = c + 1; i < n; i++)
            if (std::fabs(T[i][c]) > std::fabs(T[p][c])) p = i;
        if (T[p][c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j <= n; j++) std::swap(T[c][j], T[p][j]);
        for (int i = c + 1; i < n; i++) {
            const double f = T[i][c] / T[c][c];
            for (int j = c; j <= n; j++) T[i][j] -= f * T[c][j];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = T[i][n];
        for (int j = i + 1; j < n; j++) s -= T[i][j] * w[j];
        w[i] = s / T[i][i];
    }
    return true;
}

// io_interp_wts + io_interp: value at each xo of the data on xi, from the
// up to 7 nearest points, both grids mapped onto [0, 1].
struct Interp {
    std::vector<int> lo, cnt;
    std::vector<double> w;  // [no][7]
};

bool interp_weights(const std::vector<double>& xi, const std::vector<double>& xo, Interp& ip) {
    const int ni = (int)xi.size(), no = (int)xo.size();
    std::vector<double> x1(ni), y1(no);
    for (int k = 0; k < ni; k++) x1[k] = (xi[k] - xi[0]) / (xi[ni - 1] - xi[0]);
    for (int k = 0; k < no; k++) y1[k] = (xo[k] - xo[0]) / (xo[no - 1] - xo[0]);
    ip.lo.assign(no, 0);
    ip.cnt.assign(no, 0);
    ip.w.assign((size_t)no * 7, 0.0);
    for (int n = 0; n < no; n++) {
        int j = 0;
        while (j < ni - 1 && x1[j] < y1[n] - 1e-8) j++;
        const int l = std::max(0, j - LSD_KL), r = std::min(j + LSD_KL, ni - 1), nn = r - l + 1;
        double A[49] = {};
        for (int i = 0; i < nn; i++) A[i] = 1.0;
        for (int c = 1; c < nn; c++)
            for (int i = 0; i < nn; i++) A[c * 7 + i] = A[(c - 1) * 7 + i] * (x1[l + i] - y1[n]) / (double)c;
        ip.lo[n] = l;
        ip.cnt[n] = nn;
        if (!first_row_of_inverse(nn, A, &ip.w[(size_t)n * 7])) return false;
    }
    return true;
}

std::vector<double> radial_grid(const lsd_config& cfg) {
    const int N = cfg.N;
    const double PI = 3.1415926535897931, ri = cfg.rratio / (1.0 - cfg.rratio);
    std::vector<double> r(N);
    for (int n = 1; n <= N; n++) r[n - 1] = 0.5 * (1.0 + std::cos(PI * (N - n) / (double)(N - 1))) + ri;
    return r;
}

// ---- writer

struct Writer {
    std::vector<unsigned char> b;
    void u32(unsigned long long v) {
        for (int k = 3; k >= 0; k--) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void f64(double d) {
        unsigned long long v;
        std::memcpy(&v, &d, 8);
        for (int k = 7; k >= 0; k--) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void name(const std::string& s) {
        u32(s.size());
        b.insert(b.end(), s.begin(), s.end());
        while (b.size() % 4) b.push_back(0);
    }
    void att_d(const std::string& nm, double v) {
        name(nm);
        u32(NC_DOUBLE);
        u32(1);
        f64(v);
    }
    void att_i(const std::string& nm, int v) {
        name(nm);
        u32(NC_INT);
        u32(1);
        u32((unsigned)v);
    }
    void att_c(const std::string& nm, const std::string& v) {
        name(nm);
        u32(NC_CHAR);
        u32(v.size());
        b.insert(b.end(), v.begin(), v.end());
        while (b.size() % 4) b.push_back(0);
    }
};

}  // namespace

// coll: [6][2][H][N] (uT, uP, C, Comp, BT, BP; Re then Im; reference coll layout)
int state_read(const char* path, const lsd_config& cfg, double* coll, double* t, double* dt, StateIc* ic) {
    NcFile f;
    int rc = parse(path ? path : "", f);
    if (rc) return rc;
    const int N = cfg.N;
    std::vector<size_t> shp;
    std::vector<double> rf;
    if (!get_doubles(f, "r", shp, rf) || shp.size() != 1 || rf.size() < 2) {
        set_error("state file: no radial grid 'r'");
        return LSD_E_ARG;
    }
    const std::vector<double> ro = radial_grid(cfg);
    bool into = rf.size() != (size_t)N;
    if (!into)
        for (int n = 0; n < N; n++) into = into || std::fabs(ro[n] - rf[n]) > 1e-8;
    Interp ip;
    if (into && !interp_weights(rf, ro, ip)) {
        set_error("state file: radial interpolation weights are singular");
        return LSD_E_SINGULAR;
    }
    const int Nf = (int)rf.size();
    std::vector<int> hl, hm;
    harmonics(cfg.L, cfg.M, cfg.Mp, hl, hm);
    const int H = (int)hl.size();
    const char* names[6] = {"uT", "uP", "C", "Comp", "BT", "BP"};
    const size_t HN = (size_t)H * N;
    std::memset(coll, 0, sizeof(double) * 12 * HN);
    for (int k = 0; k < 6; k++) {
        const char* nm = names[k];
        if (k == 3 && !f.vars.count("Comp")) nm = "C";  // io_load_state: composition from codensity
        if (k >= 4 && !cfg.mag) continue;              // non-magnetic runs keep B = 0
        std::vector<double> v;
        if (!get_doubles(f, nm, shp, v)) {
            set_error(std::string("state file: no variable ") + nm);
            return LSD_E_ARG;
        }
        if (shp.size() != 3 || shp[0] != 2 || shp[2] != (size_t)Nf) {
            set_error(std::string("state file: variable ") + nm + " is not (ReIm, H, r)");
            return LSD_E_ARG;
        }
        const NcVar& var = f.vars.at(nm);
        const int fL = att_int(var, "L", cfg.L), fM = att_int(var, "M", cfg.M), fMp = att_int(var, "Mp", cfg.Mp);
        const size_t Hf = shp[1];
        for (int h = 0; h < H; h++) {
            const long j = state_nh(hl[h], hm[h], fL, fM, fMp);
            if (j < 0 || (size_t)j >= Hf) continue;  // mode not in the file
            for (int part = 0; part < 2; part++) {
                const double* src = &v[((size_t)part * Hf + (size_t)j) * Nf];
                double* dst = coll + ((size_t)(2 * k + part)) * HN + (size_t)h * N;
                if (!into) {
                    std::memcpy(dst, src, sizeof(double) * N);
                } else {
                    for (int n = 0; n < N; n++) {
                        double s = 0.0;
                        for (int c = 0; c < ip.cnt[n]; c++) s += ip.w[(size_t)n * 7 + c] * src[ip.lo[n] + c];
                        dst[n] = s;
                    }
                }
            }
        }
    }
    if (ic) {
        // icOmega, and icBT / icBP when the file has an inner core of more
        // than one point: rows 1..Nic(file) of each (ReIm, H) column of the
        // outer core's length, regridded when the radii differ
        // (io_load_coll with the inner-core grid, program/io.F90:440-447)
        std::vector<double> om, icr;
        ic->omega = get_doubles(f, "icOmega", shp, om) && om.size() == 1 ? om[0] : 0.0;
        ic->loaded = false;
        const bool has = ic->Nic > 1 && cfg.mag && f.vars.count("icBT") && f.vars.count("icBP") &&
                         get_doubles(f, "icr", shp, icr) && icr.size() > 1;
        if (has) {
            const int Ni = ic->Nic, Nif = (int)icr.size();
            bool inti = Nif != Ni;
            if (!inti)
                for (int n = 0; n < Ni; n++) inti = inti || std::fabs(icr[n] - ic->ric[n]) > 1e-8;
            Interp ipi;
            if (inti && !interp_weights(icr, ic->ric, ipi)) {
                set_error("state file: inner-core interpolation weights are singular");
                return LSD_E_SINGULAR;
            }
            ic->f.assign((size_t)4 * H * Ni, 0.0);
            const char* icn[2] = {"icBT", "icBP"};
            for (int k = 0; k < 2; k++) {
                std::vector<double> v;
                if (!get_doubles(f, icn[k], shp, v) || shp.size() != 3 || shp[0] != 2 || shp[2] < (size_t)Nif) {
                    set_error(std::string("state file: variable ") + icn[k] + " is not (ReIm, H, r)");
                    return LSD_E_ARG;
                }
                const NcVar& var = f.vars.at(icn[k]);
                const int fL = att_int(var, "L", cfg.L), fM = att_int(var, "M", cfg.M), fMp = att_int(var, "Mp", cfg.Mp);
                const size_t Hf = shp[1], Rf = shp[2];
                for (int h = 0; h < H; h++) {
                    const long j = state_nh(hl[h], hm[h], fL, fM, fMp);
                    if (j < 0 || (size_t)j >= Hf) continue;
                    for (int part = 0; part < 2; part++) {
                        const double* src = &v[((size_t)part * Hf + (size_t)j) * Rf];
                        double* dst = ic->f.data() + ((size_t)(2 * k + part) * H + h) * Ni;
                        for (int n = 0; n < Ni; n++) {
                            if (!inti) {
                                dst[n] = src[n];
                            } else {
                                double s = 0.0;
                                for (int c = 0; c < ipi.cnt[n]; c++) s += ipi.w[(size_t)n * 7 + c] * src[ipi.lo[n] + c];
                                dst[n] = s;
                            }
                        }
                    }
                }
            }
            ic->loaded = true;
        }
    }
    if (t) {
        auto it = f.gatts.find("t");
        *t = (it != f.gatts.end() && !it->second.num.empty()) ? it->second.num[0] : 0.0;
    }
    if (dt) {
        std::vector<double> d;
        *dt = get_doubles(f, "dt", shp, d) && d.size() == 1 ? d[0] : 0.0;
    }
    return LSD_OK;
}

int state_write(const char* path, const lsd_config& cfg, const double* coll, double t, double dt,
                const StateIc* ic) {
    const bool icf = ic && ic->Nic > 1 && (int)ic->ric.size() == ic->Nic;
    const int Nic = icf ? ic->Nic : 1;
    const int N = cfg.N;
    std::vector<int> hl, hm;
    harmonics(cfg.L, cfg.M, cfg.Mp, hl, hm);
    const int H = (int)hl.size();
    const size_t HN = (size_t)H * N;
    // io_state_nh index of each of this run's harmonics
    std::vector<long> pos(H);
    for (int h = 0; h < H; h++) pos[h] = state_nh(hl[h], hm[h], cfg.L, cfg.M, cfg.Mp);
    Writer w;
    w.b = {'C', 'D', 'F', 1};
    w.u32(0);  // numrecs
    // dimensions r, icr, H, ReIm
    w.u32(NC_DIMENSION);
    w.u32(4);
    w.name("r");
    w.u32(N);
    w.name("icr");
    w.u32(Nic);
    w.name("H");
    w.u32(H);
    w.name("ReIm");
    w.u32(2);
    // global attributes (io_save_state)
    w.u32(NC_ATTRIBUTE);
    w.u32(11);
    w.att_d("t", t);
    w.att_d("Ro", cfg.E / cfg.Pm);
    w.att_d("E", cfg.E);
    w.att_d("Ra", cfg.Ra);
    w.att_d("Ra_comp", cfg.Ra_comp);
    w.att_d("Pr", cfg.Pr);
    w.att_d("Sc", cfg.Sc);
    w.att_d("Pm", cfg.Pm);
    w.att_d("q", cfg.Pm / cfg.Pr);
    w.att_d("riro", cfg.rratio);
    w.att_c("LSD_version", "lsd-mi355x");
    // variables: r, icr, dt, icOmega, then the coll fields (ReIm, H, r)
    const char* names[8] = {"uT", "uP", "C", "Comp", "BT", "BP", "icBT", "icBP"};
    struct Def {
        std::string nm;
        std::vector<int> dims;
        bool coll;
        size_t count;
    };
    std::vector<Def> defs = {{"r", {0}, false, (size_t)N}, {"icr", {1}, false, (size_t)Nic}, {"dt", {}, false, 1},
                             {"icOmega", {}, false, 1}};
    for (const char* nm : names) defs.push_back({nm, {3, 2, 0}, true, 2 * HN});
    w.u32(NC_VARIABLE);
    w.u32(defs.size());
    std::vector<size_t> begin_at(defs.size());
    for (size_t i = 0; i < defs.size(); i++) {
        const Def& d = defs[i];
        w.name(d.nm);
        w.u32(d.dims.size());
        for (int x : d.dims) w.u32(x);
        if (d.coll) {
            w.u32(NC_ATTRIBUTE);
            w.u32(3);
            w.att_i("L", cfg.L);
            w.att_i("M", cfg.M);
            w.att_i("Mp", cfg.Mp);
        } else {
            w.u32(0);
            w.u32(0);
        }
        w.u32(NC_DOUBLE);
        const size_t vsize = (8 * d.count + 3) & ~(size_t)3;
        w.u32(vsize);
        begin_at[i] = w.b.size();
        w.u32(0);  // begin, patched below
    }
    size_t off = w.b.size();
    std::vector<size_t> begins(defs.size());
    for (size_t i = 0; i < defs.size(); i++) {
        begins[i] = off;
        off += 8 * defs[i].count;
    }
    if (off > 0xffffffffull) {
        set_error("lsd_write_state: state too large for the CDF-1 format");
        return LSD_E_ARG;
    }
    for (size_t i = 0; i < defs.size(); i++)
        for (int k = 0; k < 4; k++) w.b[begin_at[i] + k] = (unsigned char)(begins[i] >> (8 * (3 - k)));
    const std::vector<double> r = radial_grid(cfg);
    for (int n = 0; n < N; n++) w.f64(r[n]);
    // icr: the inner-core radii (one point, r_i, for an insulating inner core,
    // modules/meshs.F90:165-169 with i_Nic = 1)
    if (icf)
        for (int n = 0; n < Nic; n++) w.f64(ic->ric[n]);
    else
        w.f64(cfg.rratio / (1.0 - cfg.rratio));
    w.f64(dt);
    w.f64(ic ? ic->omega : 0.0);
    std::vector<double> buf((size_t)2 * HN);
    for (int k = 0; k < 8; k++) {
        std::fill(buf.begin(), buf.end(), 0.0);
        if (k < 6)
            for (int part = 0; part < 2; part++)
                for (int h = 0; h < H; h++)
                    std::memcpy(&buf[((size_t)part * H + pos[h]) * N], coll + ((size_t)(2 * k + part)) * HN + (size_t)h * N,
                                sizeof(double) * N);
        else if (icf)
            // rows 1..Nic of each column (io_save_coll with i_Nic rows)
            for (int part = 0; part < 2; part++)
                for (int h = 0; h < H; h++)
                    std::memcpy(&buf[((size_t)part * H + pos[h]) * N],
                                ic->f.data() + ((size_t)(2 * (k - 6) + part) * H + h) * Nic, sizeof(double) * Nic);
        for (double x : buf) w.f64(x);
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
        set_error(std::string("lsd_write_state: cannot open ") + path);
        return LSD_E_ARG;
    }
    const size_t put = std::fwrite(w.b.data(), 1, w.b.size(), fp);
    const int cl = std::fclose(fp);
    if (put != w.b.size() || cl != 0) {
        set_error(std::string("lsd_write_state: write failed for ") + path);
        return LSD_E_ARG;
    }
    return LSD_OK;
}

}  // namespace lsd
