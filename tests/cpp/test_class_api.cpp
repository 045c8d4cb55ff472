// C++ class layer (include/tk/*.hxx) used the way the reference's examples use
// im3d::image3d, conv::filtering, lapl::unsteady_poisson_functor and
// segm::rsfe_splitbregman (examples/rsfe_splitbregman.cxx:79-113, test.cxx), each result
// checked against the CPU oracle (oracle/liboracle.so, linked as the checker).
// Built as a shared library and called in-process by tests/test_gpu_cpp_api.py:
// tk_cpp_class_api_test(log, cap) writes one "ok <case> ..." or "FAIL <case> ..." line per
// case into log and returns 0 when every case passed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/tk/interface.hxx"
#include "../../include/tk/rsfe_splitbregman.hxx"
#include "../../oracle/rsfe_oracle.h"

static int g_fail = 0;
static std::string g_log;

static void report(const char* name, double err, double tol)
{
    const bool ok = err <= tol;
    char line[256];
    std::snprintf(line, sizeof line, "%s %s err=%.3e tol=%.1e\n", ok ? "ok" : "FAIL", name, err,
                  tol);
    g_log += line;
    if (!ok) g_fail = 1;
}

static double maxabs(const std::vector<double>& a, const std::vector<double>& b)
{
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return a.size() == b.size() ? m : INFINITY;
}

static void run_cases()
{
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> U(0., 1.);
    const unsigned X = 20, Y = 18, Z = 24;
    const long n = (long)X * Y * Z;
    std::vector<double> raw(n);
    for (auto& x : raw) x = U(gen);

    // ---- image3d: storage, copies, norms, range ----------------------------------------
    im3d::image3d<double> a(X, Y, Z, 1., 2., 0.5);
    a.upload(raw);
    report("image3d upload/download", maxabs(a.download(), raw), 0);
    im3d::image3d<double> b(a);
    report("image3d deep copy", maxabs(b.download(), raw), 0);
    b = 3.0;
    report("image3d operator=(scalar) leaves the copy source", maxabs(a.download(), raw), 0);
    report("image3d element read", std::fabs(a(3, 4, 5) - raw[(3 * Y + 4) * Z + 5]), 0);
    report("image3d normL2", std::fabs(a.normL2() - or_norm_l2(raw.data(), n, 1., 2., .5)),
           1e-12 * a.normL2());
    report("image3d max/min", std::fabs(a.max() - or_max(raw.data(), n)) +
                                  std::fabs(a.min() - or_min(raw.data(), n)), 0);
    {
        im3d::image3d<double> c(a);
        c.change_range_of_intensity(255., 0.);
        std::vector<double> ref(raw);
        or_change_range_of_intensity(ref.data(), n, 255., 0.);
        report("image3d change_range_of_intensity", maxabs(c.download(), ref), 1e-12);
    }

    // ---- grad and div ----------------------------------------------------------------
    {
        std::vector<im3d::image3d<double> > g;
        a.grad(g);
        std::vector<double> g0(n), g1(n), g2(n), dv(n);
        or_grad(raw.data(), X, Y, Z, 1., 2., .5, g0.data(), g1.data(), g2.data());
        report("image3d grad", std::fmax(maxabs(g[0].download(), g0),
                                         std::fmax(maxabs(g[1].download(), g1),
                                                   maxabs(g[2].download(), g2))), 1e-12);
        im3d::image3d<double> d;
        im3d::div(d, g);
        or_div(g0.data(), g1.data(), g2.data(), X, Y, Z, 1., 2., .5, dv.data());
        report("im3d::div", maxabs(d.download(), dv), 1e-11);
    }

    // ---- conv::filtering + cyclic_fftw_convolution --------------------------------------
    {
        im3d::image3d<double> f(X, Y, Z);
        f.upload(raw);
        conv::filtering<double> cv(conv::cyclic_fftw_convolution<double>);
        conv::filtering<double>::build_gaussian_filter(f, 2.0);
        im3d::image3d<double> out;
        cv(out, f);
        int nh[3], nc[3];
        or_gaussian_dims(Z, 1., 1., 1., 2.0, 3., 0, nh);
        std::vector<double> tf((long)nh[0] * nh[1] * nh[2]);
        or_gaussian_time_filter(Z, 1., 1., 1., 2.0, 3., 0, tf.data());
        or_cyclic_dims(X, Y, Z, nh, nc);
        std::vector<double> spec((long)nc[0] * nc[1] * nc[2]), ref(n);
        or_cyclic_spectrum(tf.data(), nh, nc, spec.data());
        or_cyclic_conv(spec.data(), nc, raw.data(), X, Y, Z, 1., 1., 1., ref.data());
        report("conv::filtering cyclic_fftw_convolution", maxabs(out.download(), ref), 1e-13);
        report("convolution result dims", (out.getdimx() == X && out.getdimy() == Y &&
                                           out.getdimz() == Z) ? 0 : 1, 0);

        // acyclic_fftw_convolution: the oracle's padded-FFT form
        cv.changef(conv::acyclic_fftw_convolution<double>);
        im3d::image3d<double> outa;
        cv(outa, f);
        int na[3];
        or_acyclic_dims(X, Y, Z, nh, na);
        std::vector<double> speca((long)na[0] * na[1] * na[2]), refa(n);
        or_cyclic_spectrum(tf.data(), nh, na, speca.data());
        or_acyclic_conv(speca.data(), na, raw.data(), X, Y, Z, 1., 1., 1., refa.data());
        report("conv::filtering acyclic_fftw_convolution", maxabs(outa.download(), refa), 1e-13);
    }

    // ---- lapl::unsteady_poisson_functor(DirGaussSeidel): bitwise --------------------------
    {
        std::vector<double> rhs(n), p0(n);
        for (auto& x : rhs) x = U(gen) - 0.5;
        for (auto& x : p0) x = U(gen);
        im3d::image3d<double> phi(X, Y, Z, 1., 2., .5), r(X, Y, Z, 1., 2., .5);
        phi.upload(p0);
        r.upload(rhs);
        lapl::unsteady_poisson_functor<double> osl(lapl::DirGaussSeidel<double>);
        for (int s = 0; s < 3; ++s) {
            osl(phi, r, 1., 0.25);
            or_dir_gauss_seidel(p0.data(), rhs.data(), X, Y, Z, 1., 2., .5, 1., 0.25);
        }
        report("unsteady_poisson_functor DirGaussSeidel x3 (bitwise)", maxabs(phi.download(), p0),
               0);

        // NeuGaussSeidel with a non-zero derivative
        std::vector<double> q0(n);
        for (auto& x : q0) x = U(gen);
        im3d::image3d<double> psi(X, Y, Z, 1., 2., .5);
        psi.upload(q0);
        osl.changef(lapl::NeuGaussSeidel<double>);
        for (int s = 0; s < 3; ++s) {
            osl(psi, r, 0.8, 0.3);
            or_neu_gauss_seidel(q0.data(), rhs.data(), X, Y, Z, 1., 2., .5, 0.8, 0.3);
        }
        report("unsteady_poisson_functor NeuGaussSeidel x3 (bitwise)", maxabs(psi.download(), q0),
               0);
    }

    // ---- connected component ----------------------------------------------------------
    {
        im3d::image3d<double> img(X, Y, Z);
        img.upload(raw);
        im3d::image3d<unsigned char> cc;
        img.connected_component(cc, 10, 9, 12, 0.5, false);
        std::vector<double> ref(n);
        or_connected_component(raw.data(), X, Y, Z, 10, 9, 12, 0.5, 0, ref.data());
        std::vector<unsigned char> got = cc.download();
        double bad = 0;
        for (long v = 0; v < n; ++v) bad += (double)got[v] != ref[v];
        report("image3d connected_component (voxels differing)", bad, 0);
    }

    // ---- segm::rsfe_splitbregman: apply(myim, init) -------------------------------------
    {
        const int SX = 24, SY = 20, SZ = 28;
        const long m = (long)SX * SY * SZ;
        std::vector<double> u0(m), init(m);
        tk_gen_standard(SX, SY, SZ, 5, 1, 2, u0.data(), init.data());
        im3d::image3d<double> myim(SX, SY, SZ), phi0(SX, SY, SZ);
        myim.upload(u0);
        phi0.upload(init);
        myim.change_range_of_intensity(255.);  // examples/rsfe_splitbregman.cxx:86
        segm::rsfe_splitbregman<double> rsfe;
        rsfe.set_verbosity(false);
        rsfe.set_cv(conv::filtering<double>(conv::cyclic_fftw_convolution<double>));
        rsfe.set_sigma(2.0);
        rsfe.set_maxiter(6);
        rsfe.set_tol(0.);
        rsfe.apply(myim, phi0);
        std::vector<double> got = rsfe.getoutput().download();

        or_params p;
        or_default_params(&p);
        p.sigma = 2.0;
        p.maxiter = 6;
        p.tol = 0.;
        std::vector<double> u255(u0), ref(m);
        or_change_range_of_intensity(u255.data(), m, 255., 0.);
        std::vector<int> sw(6);
        or_rsfe_apply(u255.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p, ref.data(), sw.data(),
                      nullptr, nullptr);
        report("rsfe_splitbregman apply FP64", maxabs(got, ref), 1e-9);
        int sweeps_equal = rsfe.get_iterations() == 6;
        for (unsigned t = 0; t < rsfe.get_iterations() && t < 6; ++t)
            sweeps_equal &= rsfe.get_sweeps()[t] == sw[t];
        report("rsfe_splitbregman sweep counts", sweeps_equal ? 0 : 1, 0);

        // apply(myim): cube initial contour, the default cv (acyclic_fftw_convolution)
        segm::rsfe_splitbregman<double> cube;
        cube.set_verbosity(false);
        cube.set_sigma(2.0);
        cube.set_maxiter(3);
        cube.set_tol(0.);
        cube.apply(myim);
        std::vector<double> cinit(m), cref(m);
        or_init_phi(nullptr, 0, 0, 0, SX, SY, SZ, 0., 1., cinit.data());
        p.maxiter = 3;
        p.acyclic = 1;
        or_rsfe_apply(u255.data(), cinit.data(), SX, SY, SZ, 1., 1., 1., &p, cref.data(),
                      nullptr, nullptr, nullptr);
        report("rsfe_splitbregman apply(myim) cube FP64", maxabs(cube.getoutput().download(), cref),
               1e-9);

        // set_bc(NEU): NeuGaussSeidel with derivative 0 (imp:1294-1322, 1736-1739)
        segm::rsfe_splitbregman<double> neu;
        neu.set_verbosity(false);
        neu.set_cv(conv::filtering<double>(conv::cyclic_fftw_convolution<double>));
        neu.set_bc(segm::rsfe_splitbregman<double>::NEU);
        neu.set_sigma(2.0);
        neu.set_maxiter(4);
        neu.set_tol(0.);
        neu.apply(myim, phi0);
        std::vector<double> nref(m);
        p.maxiter = 4;
        p.acyclic = 0;
        p.neumann = 1;
        or_rsfe_apply(u255.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p, nref.data(), nullptr,
                      nullptr, nullptr);
        report("rsfe_splitbregman set_bc(NEU) FP64", maxabs(neu.getoutput().download(), nref),
               1e-9);
        p.neumann = 0;

        // setter validation keeps the old value, as the reference
        segm::rsfe_splitbregman<double> v;
        v.set_sigma(-1.);
        v.set_b0(-2.);  // a0 moves to b0 - 1
        report("setter validation", std::fabs(v.params().sigma - 3.) +
                                        std::fabs(v.params().a0 + 3.), 0);
    }

    // ---- im3d::interface: MetaImage write / read ---------------------------------------
    char dir_t[] = "/tmp/tk_cpp_io_XXXXXX";
    const char* dir = mkdtemp(dir_t);
    if (!dir) {
        report("mkdtemp", 1, 0);
        return;
    }
    const std::string base = std::string(dir) + "/img";
    {
        im3d::image3d<double> img(X, Y, Z, 0.5, 1.25, 2.);
        img.upload(raw);
        im3d::interface<double> out(img);
        out.write(base);
        im3d::interface<double> in(base + ".mhd");
        im3d::image3d<double> back;
        in.convert2image3d(back);
        report("interface write/read (.mhd) round trip", maxabs(back.download(), raw), 0);
        report("interface spacing and dims", (back.gethx() == 0.5 && back.gethy() == 1.25 &&
                                              back.gethz() == 2. && back.getdimx() == X &&
                                              back.getdimy() == Y && back.getdimz() == Z) ? 0 : 1,
               0);
        // the .raw holds doubles with i fastest (VTK's order)
        std::vector<double> file(n);
        FILE* f = std::fopen((base + ".raw").c_str(), "rb");
        size_t got = f ? std::fread(file.data(), sizeof(double), n, f) : 0;
        if (f) std::fclose(f);
        double bad = got == (size_t)n ? 0 : 1;
        for (unsigned k = 0; k < Z && got == (size_t)n; ++k)
            for (unsigned j = 0; j < Y; ++j)
                for (unsigned i = 0; i < X; ++i)
                    bad += file[i + (long)X * (j + (long)Y * k)] != raw[((long)i * Y + j) * Z + k];
        report("interface .raw voxel order (x fastest)", bad, 0);
    }

    // ---- rsfe_splitbregman: automatic connected component and dumps ---------------------
    {
        // two bright balls; the initial contour covers both, the seed lies in the larger
        const int SX = 20, SY = 18, SZ = 16;
        const long m = (long)SX * SY * SZ;
        std::vector<double> u0(m), init(m);
        for (int i = 0; i < SX; ++i)
            for (int j = 0; j < SY; ++j)
                for (int k = 0; k < SZ; ++k) {
                    const double r1 = (i - 6) * (i - 6) + (j - 6) * (j - 6) + (k - 8) * (k - 8);
                    const double r2 = (i - 14) * (i - 14) + (j - 12) * (j - 12) + (k - 8) * (k - 8);
                    const long v = ((long)i * SY + j) * SZ + k;
                    u0[v] = 255. * (0.2 + 0.6 * (r1 <= 16) + 0.5 * (r2 <= 9)) + 10. * (U(gen) - 0.5);
                    init[v] = (r1 <= 25 || r2 <= 16) ? 1. : 0.;
                }
        im3d::image3d<double> myim(SX, SY, SZ), phi0(SX, SY, SZ);
        myim.upload(u0);
        phi0.upload(init);
        segm::rsfe_splitbregman<double> cc;
        cc.set_verbosity(false);
        cc.set_cv(conv::filtering<double>(conv::cyclic_fftw_convolution<double>));
        cc.set_sigma(1.5);
        cc.set_ls_steps(3);
        cc.set_maxiter(4);
        cc.set_tol(0.);
        cc.set_auto_extract_conn_comp(true);
        cc.set_auto_extract_conn_comp_freq(2);
        cc.set_cc_init_pixel_x(6.5);
        cc.set_cc_init_pixel_y(6.5);
        cc.set_cc_init_pixel_z(8.5);
        cc.set_dumpfrequency(3);
        cc.set_outputname(std::string(dir) + "/res");
        cc.apply(myim, phi0);
        or_params p;
        or_default_params(&p);
        p.sigma = 1.5;
        p.ls_steps = 3;
        p.maxiter = 4;
        p.tol = 0.;
        p.auto_cc = 1;
        p.auto_cc_freq = 2;
        p.cc_pixel[0] = 6.5;
        p.cc_pixel[1] = 6.5;
        p.cc_pixel[2] = 8.5;
        std::vector<double> ref(m);
        or_stats st;
        or_rsfe_apply(u0.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p, ref.data(), nullptr,
                      nullptr, &st);
        std::vector<double> got = cc.getoutput().download();
        report("rsfe_splitbregman auto connected component FP64", maxabs(got, ref), 1e-9);
        report("auto connected component: extractions", st.cc_extractions == 2 ? 0 : 1, 0);
        // dumps: res_3 (3 % 3 == 0, not maxiter) and res_4_final, no res_4
        im3d::interface<double> fin(std::string(dir) + "/res_4_final.mhd");
        im3d::image3d<double> fin_im;
        fin.convert2image3d(fin_im);
        report("dump <outputname>_4_final.mhd equals the result", maxabs(fin_im.download(), got), 0);
        const bool has3 = access((std::string(dir) + "/res_3.mhd").c_str(), F_OK) == 0 &&
                          access((std::string(dir) + "/res_3.raw").c_str(), F_OK) == 0;
        const bool no4 = access((std::string(dir) + "/res_4.mhd").c_str(), F_OK) != 0;
        report("dumps every dumpfrequency iterations, not at maxiter", (has3 && no4) ? 0 : 1, 0);
    }
    // ---- rsfe_splitbregman: parameters from a GetPot file, blurred edge detector ---------
    {
        const std::string pot = std::string(dir) + "/params.pot";
        FILE* f = std::fopen(pot.c_str(), "w");
        if (f) {
            std::fputs("# parameters of the run\n[splitbregman]\n  [./init]\n"
                       "    verbosity = 0\n    maxiter = 3   # three outer iterations\n"
                       "    sigma = 1.5\n    ls_steps=3\n    tol = 0\n"
                       "    edge_detector_sigma = 1.2\n    beta = 50\n  [../]\n"
                       "  [./onthego]\n    maxiter = 99\n  [../]\n[../]\n", f);
            std::fclose(f);
        }
        const int SX = 18, SY = 16, SZ = 20;
        const long m = (long)SX * SY * SZ;
        std::vector<double> u0(m), init(m);
        tk_gen_standard(SX, SY, SZ, 9, 1, 2, u0.data(), init.data());
        im3d::image3d<double> myim(SX, SY, SZ), phi0(SX, SY, SZ);
        myim.upload(u0);
        phi0.upload(init);
        myim.change_range_of_intensity(255.);
        segm::rsfe_splitbregman<double> gp;
        gp.set_cv(conv::filtering<double>(conv::cyclic_fftw_convolution<double>));
        gp.set_getpotfile(pot);
        gp.apply(myim, phi0);
        or_params p;
        or_default_params(&p);
        p.maxiter = 3;
        p.sigma = 1.5;
        p.ls_steps = 3;
        p.tol = 0.;
        p.edge_detector_sigma = 1.2;
        p.beta = 50;
        std::vector<double> u255(u0), ref(m);
        or_change_range_of_intensity(u255.data(), m, 255., 0.);
        or_rsfe_apply(u255.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p, ref.data(), nullptr,
                      nullptr, nullptr);
        report("rsfe_splitbregman from a GetPot file, blurred edge detector FP64",
               maxabs(gp.getoutput().download(), ref), 1e-9);
        report("GetPot: the file's maxiter (init section, not onthego)",
               gp.get_iterations() == 3 ? 0 : 1, 0);
        std::remove(pot.c_str());
    }
    for (const char* f : {"img.mhd", "img.raw", "res_3.mhd", "res_3.raw", "res_4_final.mhd",
                          "res_4_final.raw"})
        std::remove((std::string(dir) + "/" + f).c_str());
    rmdir(dir);
}

extern "C" int tk_cpp_class_api_test(char* log, int cap)
{
    g_fail = 0;
    g_log.clear();
    try {
        run_cases();
    } catch (const std::exception& e) {
        g_log += std::string("FAIL exception: ") + e.what() + "\n";
        g_fail = 1;
    }
    g_log += g_fail ? "FAILED\n" : "ALL OK\n";
    if (log && cap > 0) {
        std::strncpy(log, g_log.c_str(), (size_t)cap - 1);
        log[cap - 1] = 0;
    }
    return g_fail;
}
