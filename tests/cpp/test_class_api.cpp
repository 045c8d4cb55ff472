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

#include <sstream>

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

// user plug-ins: functions a caller of the reference could install with set_cv and
// set_onestep_poisson; apply then runs stage by stage and calls them
static int g_user_cv_calls = 0, g_user_gs_calls = 0;
static void user_cyclic(im3d::image3d<double>& r, im3d::image3d<double> const& f)
{
    ++g_user_cv_calls;
    conv::cyclic_fftw_convolution<double>(r, f);
}
static void user_gs(im3d::image3d<double>& res, im3d::image3d<double> const& b, double const& dt,
                    double const& bc, im3d::image3d<double> const& input)
{
    ++g_user_gs_calls;
    lapl::DirGaussSeidel<double>(res, b, dt, bc, input);
}

// examples/test.cxx:111-201 and 224-252 through the class layer
static void image3d_operator_cases()
{
    im3d::image3d<double> image1, image2(3, 3, 3), image3(image2);
    image1.setdim(3, 3, 3, 1.5);
    auto all = [](im3d::image3d<double> const& im, double v) {
        double e = 0;
        for (double x : im.download()) e = std::fmax(e, std::fabs(x - v));
        return e;
    };
    image2 = image1;
    image3 = 3.;
    double e = all(image1, 1.5) + all(image2, 1.5) + all(image3, 3.);
    image1 += image2;
    image3 += 1.5;
    e += all(image1, 3.) + all(image3, 4.5);
    image1 -= image2;
    image3 -= 1.5;
    e += all(image1, 1.5) + all(image3, 3.);
    image1 *= image2;
    image3 *= 1.5;
    e += all(image1, 2.25) + all(image3, 4.5);
    image1 /= image2;
    image3 /= 1.5;
    e += all(image1, 1.5) + all(image3, 3.) + all(image2, 1.5);
    report("image3d compound operators (test.cxx:119-167)", e, 0);
    e = 0;
    image3 = image1 + image2;
    e += all(image3, 3.);
    image3 = image1 + 1.5;
    e += all(image3, 3.);
    image3 = image1 - image2;
    e += all(image3, 0.);
    image3 = image1 * 1.5;
    e += all(image3, 2.25);
    image3 = image1 / image2;
    e += all(image3, 1.);
    report("image3d binary operators (test.cxx:170-195)", e + all(image1, 1.5), 0);
    image3 = 1.;
    report("scalarprod / scalarprod_L2 (test.cxx:198-202)",
           std::fabs(im3d::scalarprod(image1, image3) - 40.5) +
               std::fabs(im3d::scalarprod_L2(image1, image3) - 40.5), 0);
    // operator() write and read (test.cxx:101-107)
    image1(1, 1, 1) = 2.75;
    double a(image1(1, 1, 1));
    image1(2, 0, 1) += 1.0;
    std::vector<double> want(27, 1.5);
    want[(1 * 3 + 1) * 3 + 1] = 2.75;
    want[(2 * 3 + 0) * 3 + 1] = 2.5;
    report("image3d element write", std::fabs(a - 2.75) + maxabs(image1.download(), want), 0);
    // mismatched dims: WARNING, nothing done (image3d_imp.hxx:220-223); spacing for _L2
    im3d::image3d<double> other(3, 3, 4, 1., 1., 1.);
    other = 7.;
    im3d::image3d<double> keep(image2);
    keep += other;
    im3d::image3d<double> spaced(3, 3, 3, 2., 1., 1.);
    report("mismatched dims / spacing leave the image and give 0",
           all(keep, 1.5) + std::fabs(im3d::scalarprod(image2, other)) +
               std::fabs(im3d::scalarprod_L2(image2, spaced)), 0);
    // norms on enumerate (test.cxx:246-252): 0..26 on 3^3
    im3d::image3d<double> en(3, 3, 3);
    std::vector<double> ev(27);
    for (int i = 0; i < 27; ++i) ev[i] = i;
    en.upload(ev);
    report("norm1 / normL1 / norminf / max / min (test.cxx:246-252)",
           std::fabs(en.norm1() - 351) + std::fabs(en.normL1() - 351) +
               std::fabs(en.norminf() - 26) + std::fabs(en.max() - 26) + std::fabs(en.min()), 0);
    // norminf quirk and norm2 against the oracle on signed data
    std::mt19937_64 gen(99);
    std::normal_distribution<double> N(0., 3.);
    std::vector<double> sv(20 * 18 * 16);
    for (auto& x : sv) x = N(gen);
    sv[0] = -2.5;
    sv[100] = -40.;
    im3d::image3d<double> sg(20, 18, 16, 0.5, 1., 2.);
    sg.upload(sv);
    const long ns = (long)sv.size();
    report("norminf quirk, norm2, normL1 vs oracle",
           std::fabs(sg.norminf() - or_norminf(sv.data(), ns)) +
               std::fabs(sg.norm2() - std::sqrt(or_scalarprod(sv.data(), sv.data(), ns))) +
               std::fabs(sg.normL1() - or_norm1(sv.data(), ns)), 1e-11);
    // vector_abs of the gradient of enumerate: sqrt(1 + 9 + 81) everywhere inside
    std::vector<im3d::image3d<double> > grad;
    en.grad(grad);
    im3d::image3d<double> mag;
    im3d::vector_abs(mag, grad);
    report("vector_abs of grad(enumerate)", all(mag, std::sqrt(91.)), 1e-13);
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
    // ---- rsfe_splitbregman: the on-the-go re-read of the parameter file ------------------
    // (update_param_onthego, imp:545-578, every showfrequency iterations, imp:1816-1829)
    {
        const int SX = 18, SY = 16, SZ = 20;
        const long m = (long)SX * SY * SZ;
        std::vector<double> u0(m), init(m);
        tk_gen_standard(SX, SY, SZ, 9, 1, 2, u0.data(), init.data());
        im3d::image3d<double> myim(SX, SY, SZ), phi0(SX, SY, SZ);
        myim.upload(u0);
        phi0.upload(init);
        myim.change_range_of_intensity(255.);
        std::vector<double> u255(u0);
        or_change_range_of_intensity(u255.data(), m, 255., 0.);
        or_params p1;
        or_default_params(&p1);
        p1.maxiter = 5;
        p1.sigma = 1.5;
        p1.ls_steps = 3;
        p1.tol = 0.;
        const std::string pot = std::string(dir) + "/onthego.pot";
        auto run = [&](const char* onthego_section, segm::rsfe_splitbregman<double>& gp) {
            FILE* f = std::fopen(pot.c_str(), "w");
            if (f) {
                std::fputs("[splitbregman]\n  [./init]\n    verbosity = 0\n    maxiter = 5\n"
                           "    showfrequency = 2\n    sigma = 1.5\n    ls_steps = 3\n"
                           "    tol = 0\n  [../]\n  [./onthego]\n", f);
                std::fputs(onthego_section, f);
                std::fputs("  [../]\n[../]\n", f);
                std::fclose(f);
            }
            gp.set_cv(conv::filtering<double>(conv::cyclic_fftw_convolution<double>));
            gp.set_getpotfile(pot);
            gp.apply(myim, phi0);
        };
        // after iteration 2 (and again, unchanged, after 4): sigma, lamda1, ls_steps, dt,
        // lamda change; the filter, kones and kmyim are built again
        segm::rsfe_splitbregman<double> ga;
        run("    sigma = 1.2\n    lamda1 = 1.4e-5\n    ls_steps = 5\n    dt = 0.9\n"
            "    lamda = 0.08\n", ga);
        or_params p2 = p1;
        p2.sigma = 1.2;
        p2.lamda1 = 1.4e-5;
        p2.ls_steps = 5;
        p2.dt = 0.9;
        p2.lamda = 0.08;
        std::vector<double> ref(m);
        std::vector<int> sw(5, 0);
        or_rsfe_apply_onthego(u255.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p1, 2, &p2, 0,
                              ref.data(), sw.data(), nullptr, nullptr);
        report("on the go: sigma, lamda1, ls_steps, dt, lamda from section onthego/ FP64",
               maxabs(ga.getoutput().download(), ref), 1e-9);
        report("on the go: sweeps per iteration 3 3 5 5 5",
               (ga.get_sweeps() == std::vector<int>{3, 3, 5, 5, 5} &&
                sw == std::vector<int>{3, 3, 5, 5, 5}) ? 0 : 1, 0);
        // save_current and end_now: phi written as <outputname>_2, the loop ends after 2
        segm::rsfe_splitbregman<double> gb;
        const std::string og = std::string(dir) + "/og";
        run(("    save_current = 1\n    end_now = 1\n    outputname = '" + og + "'\n").c_str(),
            gb);
        or_params p3 = p1;
        p3.maxiter = 2;
        or_rsfe_apply(u255.data(), init.data(), SX, SY, SZ, 1., 1., 1., &p3, ref.data(), nullptr,
                      nullptr, nullptr);
        report("on the go: end_now ends the loop after the re-read", gb.get_iterations() == 2
                                                                     ? 0 : 1, 0);
        report("on the go: end_now result FP64", maxabs(gb.getoutput().download(), ref), 1e-9);
        im3d::interface<double> sv(og + "_2.mhd");
        im3d::image3d<double> sv_im;
        sv.convert2image3d(sv_im);
        report("on the go: save_current writes <outputname>_2", maxabs(sv_im.download(), ref),
               1e-9);
        std::remove(pot.c_str());
    }
    image3d_operator_cases();

    // ---- conv::direct_convolution (convolution_imp.hxx:761-812) --------------------------
    {
        im3d::image3d<double> img(X, Y, Z, 1., 1., 1.), res;
        img.upload(raw);
        conv::filtering<double>::build_gaussian_filter(img, 1.7);
        conv::direct_convolution<double>(res, img);
        int nh[3];
        or_gaussian_dims(Z, 1., 1., 1., 1.7, 3., 0, nh);
        std::vector<double> tf((long)nh[0] * nh[1] * nh[2]), ref(n);
        or_gaussian_time_filter(Z, 1., 1., 1., 1.7, 3., 0, tf.data());
        or_direct_acyclic_conv(tf.data(), nh, raw.data(), X, Y, Z, ref.data());
        report("conv::direct_convolution vs the oracle's triple sum", maxabs(res.download(), ref),
               1e-13);
    }

    // ---- set_cv / set_onestep_poisson with user functions: the stage-by-stage apply -------
    {
        const int SX = 24, SY = 20, SZ = 28;
        const long m = (long)SX * SY * SZ;
        std::vector<double> u0(m), init(m);
        tk_gen_standard(SX, SY, SZ, 5, 1, 2, u0.data(), init.data());
        im3d::image3d<double> myim(SX, SY, SZ), phi0(SX, SY, SZ);
        myim.upload(u0);
        phi0.upload(init);
        myim.change_range_of_intensity(255.);
        auto run = [&](conv::filtering<double>::function cv,
                       lapl::unsteady_poisson_functor<double>::function gs, bool* fused) {
            segm::rsfe_splitbregman<double> r;
            r.set_verbosity(false);
            r.set_cv(conv::filtering<double>(cv));
            r.set_onestep_poisson(lapl::unsteady_poisson_functor<double>(gs));
            r.set_sigma(2.0);
            r.set_maxiter(4);
            r.set_tol(0.);
            *fused = r.fused();
            r.apply(myim, phi0);
            return std::make_pair(r.getoutput().download(), r.get_sweeps());
        };
        bool f0 = false, f1 = true, f2 = true, f3 = true;
        auto builtin = run(conv::cyclic_fftw_convolution<double>, lapl::DirGaussSeidel<double>, &f0);
        g_user_cv_calls = g_user_gs_calls = 0;
        auto ucv = run(user_cyclic, lapl::DirGaussSeidel<double>, &f1);
        const int cv_calls = g_user_cv_calls;
        auto ugs = run(conv::cyclic_fftw_convolution<double>, user_gs, &f2);
        const int gs_calls = g_user_gs_calls;
        report("set_cv(user function): stage-by-stage apply equals the fused loop",
               (f0 && !f1 && cv_calls == 2 + 4 * 4 && ucv.second == builtin.second)
                   ? maxabs(ucv.first, builtin.first) : 1., 1e-9);
        report("set_onestep_poisson(user function): stage-by-stage apply equals the fused loop",
               (!f2 && gs_calls == 4 * 5 && ugs.second == builtin.second)
                   ? maxabs(ugs.first, builtin.first) : 1., 1e-9);
        // conv::direct_convolution as cv: the acyclic result of the fused default cv
        auto acyc = run(conv::acyclic_fftw_convolution<double>, lapl::DirGaussSeidel<double>, &f3);
        bool f4 = true;
        auto dcv = run(conv::direct_convolution<double>, lapl::DirGaussSeidel<double>, &f4);
        report("set_cv(direct_convolution) equals the fused acyclic loop",
               (f3 && !f4) ? maxabs(dcv.first, acyc.first) : 1., 1e-9);
        // the verbose record on std::clog (imp:1655-1658, 1786-1793, 1835-1844), both paths
        std::ostringstream cap;
        std::streambuf* old = std::clog.rdbuf(cap.rdbuf());
        for (bool user : {false, true}) {
            segm::rsfe_splitbregman<double> r;
            r.set_cv(conv::filtering<double>(user ? user_cyclic : conv::cyclic_fftw_convolution<double>));
            r.set_sigma(2.0);
            r.set_maxiter(2);
            r.set_tol(0.);
            r.apply(myim, phi0);
        }
        std::clog.rdbuf(old);
        const std::string log = cap.str();
        size_t cnt = 0;
        for (size_t q = log.find("norm_inf: "); q != std::string::npos; q = log.find("norm_inf: ", q + 1))
            ++cnt;
        report("verbose record: 2 iterations x 2 paths on std::clog",
               (log.find("-- Iteration 2") != std::string::npos &&
                log.find("linear system error at 5th (last) step: ") != std::string::npos &&
                cnt == 4) ? 0 : 1, 0);
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
