// fwdfem / fwdfem_cw: the command-line forward model of the reference
// (src/supertoast/fwdfem.cc: frequency domain, complex fields;
// src/supertoast/fwdfem_cw.cc: continuous wave, real fields) on
// libtoastpp_amd.  Built twice from this file (TMI_FWDFEM_CW selects the CW
// program).
//
// Same dialogue: answers come from stdin in the order the reference asks
// for them, or from a parameter file given as the first argument
// ("KEY = value" lines, ParamParser, pparse.cc:16-125); the second argument
// names the log (default fwdfem.out), which records every choice in the
// same "KEY = value" form.  Same outputs in the working directory:
//   fwdfem:    fmod.fem + farg.fem (log data) or fre.fem + fim.fem (linear);
//              nodal fields phi_lnmod.nim + phi_arg.nim or phi_re.nim + phi_im.nim
//   fwdfem_cw: fmod.fem (log) or fre.fem; nodal fields lphi.nim or phi.nim
//
// What the library does not provide is refused with a message and exit
// status 1 before any GPU work: Tri3 and Tet4 meshes only, linear solvers
// DIRECT / CG / BICGSTAB / GMRES (no BiCG, Gauss-Seidel; CG only for the
// real CW system), preconditioners NONE / DIAG (no ICH, DILU), source and
// measurement profiles POINT / GAUSSIAN (no COSINE, no trigonometric
// sources).  DIRECT is the library's FP64-accurate iterative solve
// (relative residual 1e-14).  Non-convergence is reported and the results
// are still written, as the reference does.
#include <strings.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/toastpp_amd.h"

namespace {

#ifdef TMI_FWDFEM_CW
constexpr bool CPLX = false;
const char *PROG = "fwdfem_cw";
#else
constexpr bool CPLX = true;
const char *PROG = "fwdfem";
#endif

[[noreturn]] void die(const char *fmt, ...)
{
    std::cout << std::flush;
    fprintf(stderr, "%s: ", PROG);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    exit(1);
}

void check(int rc, const char *what)
{
    if (rc != TMI_OK) die("%s failed (status %d): %s", what, rc, tmi_last_error());
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r' || s[a] == '\n')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r' || s[b - 1] == '\n')) b--;
    return s.substr(a, b - a);
}

// Parameter file and log (ParamParser semantics, pparse.cc:16-125): a key
// matches a line that starts with it (any case) followed by ' ' or '=';
// the value is the text after '=', with a ';' comment and surrounding
// blanks removed.
struct Params {
    std::vector<std::string> lines;
    std::ofstream log;

    bool load(const char *path)
    {
        std::ifstream f(path);
        if (!f.good()) return false;
        std::string l;
        while (std::getline(f, l)) lines.push_back(l);
        return true;
    }
    void open_log(const char *path)
    {
        log.open(path);
        log << "================== TOAST LOG ==================\n" << path << "\n";
    }
    bool get(const char *key, std::string &val) const
    {
        size_t lk = strlen(key);
        for (const std::string &l : lines) {
            if (l.size() > lk && strncasecmp(l.c_str(), key, lk) == 0 && (l[lk] == ' ' || l[lk] == '=')) {
                size_t eq = l.find('=');
                if (eq == std::string::npos) return false;
                std::string v = l.substr(eq + 1);
                size_t sc = v.find(';');
                if (sc != std::string::npos) v.resize(sc);
                val = trim(v);
                return true;
            }
        }
        return false;
    }
    bool get_real(const char *key, double &v) const
    {
        std::string s;
        return get(key, s) && sscanf(s.c_str(), "%lf", &v) == 1;
    }
    bool get_int(const char *key, int &v) const
    {
        std::string s;
        return get(key, s) && sscanf(s.c_str(), "%d", &v) == 1;
    }
    void put(const char *key, const std::string &v) { log << key << " = " << v << "\n" << std::flush; }
    void put_real(const char *key, double v)
    {
        char b[64];
        snprintf(b, sizeof b, "%g", v);
        put(key, b);
    }
    void put_int(const char *key, int v) { put(key, std::to_string(v)); }
};

template <class T>
T ask(const char *prompt)
{
    std::cout << prompt << std::flush;
    T v;
    if (!(std::cin >> v)) die("input ended (or is not a valid answer) at: %s", trim(prompt).c_str());
    return v;
}

enum Solver { S_DIRECT = 1, S_CG, S_BICG, S_BICGSTAB, S_GMRES, S_GS };
const char *SOLVER_NAMES[] = {"", "DIRECT", "CG", "BICG", "BICGSTAB", "GMRES", "GAUSSSEIDEL"};
enum Precon { P_NONE = 0, P_DIAG, P_ICH, P_DILU };
const char *PRECON_NAMES[] = {"NONE", "DIAG", "ICH", "DILU"};
const char *PROFILE_NAMES[] = {"POINT", "GAUSSIAN", "COSINE"};

struct Config {
    std::string meshname, qmname;
    int dscale = -1;            // TMI_DATA_LIN / TMI_DATA_LOG
    int solver = 0;             // Solver
    double tol = 1e-10;
    int maxit = 0;
    int precon = -1;            // Precon
    bool neumann = true;
    int qprof = -1, mprof = -1;  // 0 point, 1 Gaussian, 2 cosine, 3 trigonometric
    double qwidth = 0, mwidth = 0;
    double freq = 0;
};

// Mesh::operator>> plus region numbers (node.cc:130-162)
struct HostMesh {
    int dim = 0, n = 0, ne = 0, nn = 0;
    std::vector<double> nodes;
    std::vector<int> elist, region;
};

HostMesh read_mesh(const std::string &path)
{
    HostMesh h;
    check(tmi_read_mesh_file(path.c_str(), &h.dim, &h.n, &h.ne, &h.nn, nullptr, nullptr, nullptr),
          ("reading mesh " + path).c_str());
    h.nodes.resize((size_t)h.n * h.dim);
    h.elist.resize((size_t)h.ne * h.nn);
    h.region.resize(h.n);
    check(tmi_read_mesh_file(path.c_str(), &h.dim, &h.n, &h.ne, &h.nn, h.nodes.data(), h.elist.data(),
                             h.region.data()),
          ("reading mesh " + path).c_str());
    return h;
}

// SelectMesh (fwdfem.cc:466-488): mesh and QM file names
void select_mesh(Params &pp, Config &c, HostMesh &hm)
{
    if (!pp.get("MESHFILE", c.meshname)) c.meshname = ask<std::string>("\nMesh file name:\n>> ");
    hm = read_mesh(c.meshname);
    if (!pp.get("QMFILE", c.qmname)) c.qmname = ask<std::string>("\nQM file name:\n>> ");
    pp.put("MESHFILE", c.meshname);
    pp.put("QMFILE", c.qmname);
}

// TFwdSolver::ReadParams / WriteParams (fwdsolver.cc:1084-1244)
void select_solver(Params &pp, Config &c)
{
    std::string s;
    if (pp.get("DATASCALE", s)) {
        if (!strcasecmp(s.c_str(), "LIN")) c.dscale = TMI_DATA_LIN;
        else if (!strcasecmp(s.c_str(), "LOG")) c.dscale = TMI_DATA_LOG;
    }
    while (c.dscale < 0) {
        int cmd = ask<int>("\nSelect data scaling:\n(1) Linear\n(2) Logarithmic\n[1|2] >> ");
        if (cmd == 1) c.dscale = TMI_DATA_LIN;
        else if (cmd == 2) c.dscale = TMI_DATA_LOG;
    }
    if (pp.get("LINSOLVER", s)) {
        for (int k = S_DIRECT; k <= S_GS; k++)
            if (!strcasecmp(s.c_str(), SOLVER_NAMES[k])) c.solver = k;
    }
    while (c.solver == 0) {
        int cmd = ask<int>(
            "\nSelect linear solver:\n(1) Direct\n(2) CG (Conjugate gradient)\n(3) BiCG (Bi-conjugate gradient)\n"
            "(4) BiCG-STAB (Bi-conjugate gradient stabilised)\n(5) GMRES (generalised minimum residual)\n"
            "(6) GS (Gauss-Seidel)\n>> ");
        if (cmd >= S_DIRECT && cmd <= S_GS) c.solver = cmd;
    }
    if (c.solver != S_DIRECT) {
        if (!pp.get_real("LINSOLVER_TOL", c.tol)) {
            do {
                c.tol = ask<double>("\nLinear solver stopping criterion:\n>> ");
            } while (c.tol <= 0.0);
        }
        if (!pp.get_int("LINSOLVER_MAXIT", c.maxit))
            c.maxit = ask<int>("\nLinear solver iteration limit (0=auto):\n>> ");
        if (pp.get("LINSOLVER_PRECON", s)) {
            for (int k = P_NONE; k <= P_DILU; k++)
                if (!strcasecmp(s.c_str(), PRECON_NAMES[k])) c.precon = k;
        }
        while (c.precon < 0) {
            int cmd = ask<int>(
                "\nSelect preconditioner for linear solver:\n(0) None\n(1) Diagonal\n(2) Incomplete Choleski\n"
                "(3) Diagonal incomplete LU\n");
            if (cmd >= P_NONE && cmd <= P_DILU) c.precon = cmd;
        }
    }
    pp.put("DATASCALE", c.dscale == TMI_DATA_LIN ? "LIN" : "LOG");
    pp.put("LINSOLVER", SOLVER_NAMES[c.solver]);
    if (c.solver != S_DIRECT) {
        pp.put_real("LINSOLVER_TOL", c.tol);
        pp.put_int("LINSOLVER_MAXIT", c.maxit);
        pp.put("LINSOLVER_PRECON", PRECON_NAMES[c.precon]);
    }
    if (c.solver == S_BICG || c.solver == S_GS)
        die("linear solver %s is not provided (DIRECT, CG, BICGSTAB, GMRES)", SOLVER_NAMES[c.solver]);
    if (CPLX && c.solver == S_CG)
        die("CG is not defined for the complex system (the reference's complex CG returns without solving); "
            "use BICGSTAB");
    if (c.solver != S_DIRECT && (c.precon == P_ICH || c.precon == P_DILU))
        die("preconditioner %s is not provided (NONE, DIAG)", PRECON_NAMES[c.precon]);
    if (c.maxit < 0) die("iteration limit %d is negative", c.maxit);
}

// one profile selection (SelectSourceProfile / SelectMeasurementProfile,
// fwdfem.cc:492-611): 0 point, 1 Gaussian, 2 cosine, 3 trigonometric
int select_profile(Params &pp, const char *key, const char *wkey, const char *what, double &width)
{
    std::string s;
    int t = -1;
    if (pp.get(key, s)) {
        for (int k = 0; k < 3; k++)
            if (!strcasecmp(s.c_str(), PROFILE_NAMES[k])) t = k;
    }
    char prompt[160];
    snprintf(prompt, sizeof prompt, "\n%s profile:\n(1) Point\n(2) Gaussian\n(3) Cosine\n[1|2|3] >> ", what);
    while (t < 0) t = ask<int>(prompt) - 1;
    if (t == 2 || t == 3) die("%s profile %s is not provided (POINT, GAUSSIAN)", what, t == 2 ? "COSINE" : "TRIG");
    if (t > 3) die("%s profile %d is not defined", what, t + 1);
    if (t > 0 && !pp.get_real(wkey, width)) {
        snprintf(prompt, sizeof prompt, "\n%s 1/e radius [mm]:\n>> ", what);
        width = ask<double>(prompt);
    }
    pp.put(key, PROFILE_NAMES[t]);
    if (t > 0) pp.put_real(wkey, width);
    return t;
}

void select_profiles(Params &pp, Config &c)
{
    std::string s;
    bool ok = false;
    if (pp.get("SOURCETYPE", s)) {
        if (!strcasecmp(s.c_str(), "NEUMANN")) c.neumann = ok = true;
        else if (!strcasecmp(s.c_str(), "ISOTROPIC")) { c.neumann = false; ok = true; }
    }
    while (!ok) {
        int cmd = ask<int>("\nSource type:\n(1) Neumann boundary source\n(2) Isotropic point source\n[1|2] >> ");
        if (cmd == 1) c.neumann = ok = true;
        else if (cmd == 2) { c.neumann = false; ok = true; }
    }
    pp.put("SOURCETYPE", c.neumann ? "NEUMANN" : "ISOTROPIC");
    c.qprof = select_profile(pp, "SOURCEPROFILE", "SOURCEWIDTH", "Source", c.qwidth);
    c.mprof = select_profile(pp, "MEASUREMENTPROFILE", "MEASUREMENTWIDTH", "Measurement", c.mwidth);
}

// SelectData (fwdfem.cc:615-623)
void select_freq(Params &pp, Config &c)
{
    if (!pp.get_real("FREQ", c.freq) || c.freq < 0.0) {
        do {
            c.freq = ask<double>("\nSource modulation frequency [MHz]:\n>> ");
        } while (c.freq < 0.0);
    }
    pp.put_real("FREQ", c.freq);
}

std::vector<double> read_nim_first(const std::string &path, int n)
{
    int size = 0;
    check(tmi_read_nim(path.c_str(), 0, &size, nullptr, nullptr), ("reading NIM " + path).c_str());
    if (size != n) die("NIM file %s has %d values per image, the mesh %d nodes", path.c_str(), size, n);
    std::vector<double> v(n);
    check(tmi_read_nim(path.c_str(), 0, nullptr, nullptr, v.data()), ("reading NIM " + path).c_str());
    return v;
}

// SelectInitialParams (fwdfem.cc:641-729): mua, mus, n per node from a
// global value, values per region (regions present, ascending) or the
// first image of a NIM file.
void select_params(Params &pp, const HostMesh &hm, std::vector<double> prm[3])
{
    const char *keys[3] = {"RESET_MUA", "RESET_MUS", "RESET_N"};
    const int MAXREGION = 100;
    std::vector<int> count(MAXREGION, 0);
    for (int r : hm.region)
        if (r >= 0 && r < MAXREGION) count[r]++;
    for (int p = 0; p < 3; p++) {
        std::vector<double> &v = prm[p];
        v.assign(hm.n, 0.0);
        std::string s;
        bool from_file = pp.get(keys[p], s);
        if (!from_file) {
            std::cout << "\nSelect initial distribution for " << keys[p] << "\n"
                      << "(1) Global homogeneous\n(2) Homogeneous in regions\n(3) Nodal image file (NIM)\n";
            int t = ask<int>("[1|2|3] >> ");
            char b[64];
            if (t == 1) {
                double x = ask<double>("\nGlobal value:\n>> ");
                snprintf(b, sizeof b, "HOMOG %f", x);
                s = b;
            } else if (t == 2) {
                s = "REGION_HOMOG";
                for (int r = 0; r < MAXREGION; r++) {
                    if (!count[r]) continue;
                    char pr[96];
                    snprintf(pr, sizeof pr, "Value for region %d (%d nodes):\n>> ", r, count[r]);
                    double x = ask<double>(pr);
                    snprintf(b, sizeof b, " %f", x);
                    s += b;
                }
            } else if (t == 3) {
                s = "NIM " + ask<std::string>("\nNIM file name:\n>> ");
            } else {
                die("invalid choice %d for %s", t, keys[p]);
            }
        }
        pp.put(keys[p], s);
        if (!strcasecmp(s.c_str(), "MESH")) {
            die("%s = MESH is no longer supported", keys[p]);
        } else if (!strncasecmp(s.c_str(), "REGION_HOMOG", 12)) {
            std::vector<double> vals;
            char *end = nullptr;
            const char *q = s.c_str() + 12;
            for (;;) {
                double x = strtod(q, &end);
                if (end == q) break;
                vals.push_back(x);
                q = end;
            }
            size_t k = 0;
            for (int r = 0; r < MAXREGION && k < vals.size(); r++) {
                if (!count[r]) continue;
                for (int j = 0; j < hm.n; j++)
                    if (hm.region[j] == r) v[j] = vals[k];
                k++;
            }
        } else if (!strncasecmp(s.c_str(), "HOMOG", 5)) {
            double x = 0;
            if (sscanf(s.c_str() + 5, "%lf", &x) != 1) die("%s = %s: no value", keys[p], s.c_str());
            v.assign(hm.n, x);
        } else if (!strncasecmp(s.c_str(), "NIM", 3)) {
            v = read_nim_first(trim(s.substr(3)), hm.n);
        } else {
            die("%s = %s is not understood (HOMOG, REGION_HOMOG, NIM)", keys[p], s.c_str());
        }
    }
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// dense profile rows -> CSR (row k = profile k), complex values interleaved
// when the solver is complex
void to_csr(const std::vector<double> &dense, int rows, int n, std::vector<int> &rp, std::vector<int> &ci,
            std::vector<double> &val)
{
    rp.assign(1, 0);
    ci.clear();
    val.clear();
    for (int r = 0; r < rows; r++) {
        for (int j = 0; j < n; j++) {
            double x = dense[(size_t)r * n + j];
            if (x == 0.0) continue;
            ci.push_back(j);
            val.push_back(x);
            if (CPLX) val.push_back(0.0);
        }
        rp.push_back((int)ci.size());
    }
}

// profiles of `pts` (npts x dim) as dense rows: QVec_Point / QVec_Gaussian
// (source.cc:4-41, 157-196), scaled nodally by `scale` when given
std::vector<double> profiles(tmi_mesh m, int n, int npts, const std::vector<double> &pts, int prof, double width,
                             bool neumann, const double *scale)
{
    std::vector<double> out((size_t)npts * n, 0.0);
    if (npts == 0) return out;
    if (prof == 0) {
        check(tmi_qvec_point(m, npts, pts.data(), neumann ? 1 : 0, out.data()), "point profiles");
        if (scale)
            for (int k = 0; k < npts; k++)
                for (int j = 0; j < n; j++) out[(size_t)k * n + j] *= scale[j];
    } else {
        check(tmi_qvec_gaussian(m, npts, pts.data(), width, neumann ? 1 : 0, scale, out.data()),
              "Gaussian profiles");
    }
    return out;
}

}  // namespace

int main(int argc, char **argv)
{
    Params pp;
    if (argc > 1 && pp.load(argv[1])) std::cout << "Reading parameters from " << argv[1] << std::endl;
    const char *logname = argc > 2 ? argv[2] : "fwdfem.out";
    pp.open_log(logname);
    std::cout << "Writing log to " << logname << std::endl;

    // the whole dialogue first: nothing runs on the GPU until every choice
    // is known to be supported
    Config c;
    HostMesh hm;
    select_mesh(pp, c, hm);
    select_solver(pp, c);
    select_profiles(pp, c);
    if (CPLX) select_freq(pp, c);
    std::vector<double> prm[3];
    select_params(pp, hm, prm);
    const int n = hm.n;
    const double omega = c.freq * 2.0 * M_PI * 1e-6;

    tmi_mesh mesh = nullptr;
    check(tmi_mesh_make(hm.dim, n, hm.nodes.data(), hm.ne, hm.elist.data(), &mesh), "mesh setup");
    check(tmi_mesh_read_qm(mesh, c.qmname.c_str()), ("reading QM " + c.qmname).c_str());
    int nq = 0, nm = 0, nqm = 0;
    check(tmi_mesh_qm_dims(mesh, &nq, &nm, &nqm), "QM sizes");
    std::vector<double> Q((size_t)nq * hm.dim), M((size_t)nm * hm.dim);
    std::vector<int> lptr(nq + 1), lidx(nqm > 0 ? nqm : 1);
    check(tmi_mesh_get_qm(mesh, Q.data(), M.data(), lptr.data(), lidx.data()), "QM positions");

    std::vector<double> cmua(n), ckappa(n), c2a(n);
    check(tmi_optical_to_coeffs(n, prm[0].data(), prm[1].data(), prm[2].data(), cmua.data(), ckappa.data(),
                                c2a.data()),
          "optical coefficients");

    // source and measurement vectors (fwdfem.cc:131-178): M scaled by c2a
    std::vector<double> qd = profiles(mesh, n, nq, Q, c.qprof, c.qwidth, c.neumann, nullptr);
    std::vector<double> md = profiles(mesh, n, nm, M, c.mprof, c.mwidth, true, c2a.data());
    std::vector<int> qrp, qci, mrp, mci;
    std::vector<double> qval, mval;
    to_csr(qd, nq, n, qrp, qci, qval);
    to_csr(md, nm, n, mrp, mci, mval);
    qd.clear();
    md.clear();

    int ls = c.solver == S_DIRECT ? TMI_LSOLVER_DIRECT
             : c.solver == S_CG   ? TMI_LSOLVER_CG
             : c.solver == S_GMRES ? TMI_LSOLVER_GMRES
                                   : TMI_LSOLVER_BICGSTAB;
    tmi_solver s = nullptr;
    check(tmi_solver_create(mesh, CPLX ? 1 : 0, ls, c.tol, &s), "solver setup");
    if (c.solver != S_DIRECT) {
        check(tmi_solver_set_precon(s, c.precon == P_DIAG ? TMI_PRECON_DIAG : TMI_PRECON_NONE), "preconditioner");
        check(tmi_solver_set_maxit(s, c.maxit), "iteration limit");
    }
    std::cout << "\nAllocating system matrix" << std::endl;
    check(tmi_solver_allocate(s), "allocating the system matrix");
    std::cout << "\n\n----------\nAssembling and pre-processing system matrix" << std::endl;
    double t0 = now();
    check(tmi_solver_reset(s, cmua.data(), ckappa.data(), c2a.data(), omega), "assembly");
    check(tmi_synchronize(), "assembly");
    double dt = now() - t0;
    std::cout << "T(assembly): " << dt << " real" << std::endl;
    pp.put_real("T(assembly)", dt);

    std::cout << "Computing fields" << std::endl;
    const int cw = CPLX ? 2 : 1;
    std::vector<double> phi((size_t)nq * n * cw, 0.0);
    std::vector<tmi_colstat> st(nq > 0 ? nq : 1);
    t0 = now();
    int rc = nq > 0 ? tmi_solver_calc_fields(s, nq, qrp.data(), qci.data(), qval.data(), phi.data(), st.data())
                    : TMI_OK;
    if (rc != TMI_OK && rc != TMI_ERR_NOT_CONVERGED) check(rc, "field solve");
    dt = now() - t0;
    std::cout << "T(solve): " << dt << " real" << std::endl;
    pp.put_real("T(solve)", dt);
    if (c.solver != S_DIRECT) {
        long its = 0;
        double err = 0;
        for (int q = 0; q < nq; q++) {
            its += st[q].iters;
            err = std::max(err, st[q].relres);
        }
        std::cout << "Tot. iterations: " << its << "\nMax. rel. error: " << err << std::endl;
    }
    if (rc == TMI_ERR_NOT_CONVERGED)
        std::cout << "Warning: some sources ended above the stopping criterion: " << tmi_last_error() << std::endl;

    // nodal fields (fwdfem.cc:204-231, fwdfem_cw.cc:196-216)
    if (ask<int>("Output nodal fields (1/0)? ")) {
        const char *mn = c.meshname.c_str();
        if (CPLX) {
            std::vector<double> a((size_t)nq * n), b((size_t)nq * n);
            for (size_t k = 0; k < (size_t)nq * n; k++) {
                std::complex<double> z(phi[2 * k], phi[2 * k + 1]);
                a[k] = c.dscale == TMI_DATA_LOG ? std::log(std::abs(z)) : z.real();
                b[k] = c.dscale == TMI_DATA_LOG ? std::arg(z) : z.imag();
            }
            const char *fa = c.dscale == TMI_DATA_LOG ? "phi_lnmod.nim" : "phi_re.nim";
            const char *fb = c.dscale == TMI_DATA_LOG ? "phi_arg.nim" : "phi_im.nim";
            if (nq > 0) {
                check(tmi_write_nim(fa, mn, n, nq, a.data(), TMI_NIM_FWDFEM), fa);
                check(tmi_write_nim(fb, mn, n, nq, b.data(), TMI_NIM_FWDFEM), fb);
            }
            std::cout << "  " << (c.dscale == TMI_DATA_LOG ? "Log Mod" : "Real") << " field written to " << fa
                      << "\n  " << (c.dscale == TMI_DATA_LOG ? "Arg" : "Imag") << " field written to " << fb
                      << std::endl;
        } else {
            std::vector<double> a(phi);
            if (c.dscale == TMI_DATA_LOG)
                for (double &x : a) x = std::log(x);
            const char *fa = c.dscale == TMI_DATA_LOG ? "lphi.nim" : "phi.nim";
            if (nq > 0) check(tmi_write_nim(fa, mn, n, nq, a.data(), TMI_NIM_FWDFEM), fa);
            std::cout << "  " << (c.dscale == TMI_DATA_LOG ? "Log" : "Real") << " field written to " << fa
                      << std::endl;
        }
    }

    // data files (fwdfem.cc:234-250, fwdfem_cw.cc:219-232)
    if (ask<int>("Output data files (1/0)? ")) {
        std::vector<double> proj((size_t)(nqm > 0 ? nqm : 1) * cw);
        if (nqm > 0)
            check(tmi_solver_project_all(s, nm, mrp.data(), mci.data(), mval.data(), phi.data(), c.dscale,
                                         proj.data()),
                  "projection");
        if (CPLX) {
            std::vector<double> re(nqm), im(nqm);
            for (int k = 0; k < nqm; k++) {
                re[k] = proj[2 * k];
                im[k] = proj[2 * k + 1];
            }
            const char *fa = c.dscale == TMI_DATA_LOG ? "fmod.fem" : "fre.fem";
            const char *fb = c.dscale == TMI_DATA_LOG ? "farg.fem" : "fim.fem";
            check(tmi_write_vector(fa, nqm, re.data()), fa);
            check(tmi_write_vector(fb, nqm, im.data()), fb);
        } else {
            const char *fa = c.dscale == TMI_DATA_LOG ? "fmod.fem" : "fre.fem";
            check(tmi_write_vector(fa, nqm, proj.data()), fa);
        }
    }
    tmi_solver_free(s);
    tmi_mesh_free(mesh);
    return 0;
}
