/*
 * tk/rsfe_splitbregman.hxx -- segm::segmentation<T> and segm::rsfe_splitbregman<T> over the
 * device Split-Bregman solver of libtk (tk_rsfe_*).
 *
 * Mirrors segm::segmentation<T> (header/segmentation.hxx:50-130: apply(myim) and
 * apply(myim, init) pure virtual, getoutput() returning phi by value) and
 * segm::rsfe_splitbregman<T> (header/rsfe_splitbregman.hxx:99-894): the constructor
 * defaults (rsfe_splitbregman_imp.hxx:25-41), the setters of the parameters the loop reads
 * with the reference's validation and WARNING-and-keep behaviour (imp:990-1330), set_cv and
 * set_onestep_poisson (imp:836-852), set_bc (imp:1294-1322), and apply (imp:1547-1995).
 *
 * What this layer runs.  With the built-in plug-ins -- cv = conv::acyclic_fftw_convolution
 * (the reference's default) or conv::cyclic_fftw_convolution, onestep_poisson =
 * lapl::DirGaussSeidel (bit-exact lexicographic sweeps), lapl::DirRedBlackGaussSeidel
 * (red-black) or lapl::NeuGaussSeidel (installed by set_bc(NEU), as in the reference) --
 * apply runs the fused device loop (tk_rsfe_*).  With any other function installed by set_cv
 * or set_onestep_poisson, apply runs the reference's sequence stage by stage on device
 * image3d objects and calls the installed functors, as the reference does (imp:1683-1936);
 * that path leaves out the automatic connected component and the dumps (WARNING).  Edge
 * detector GRAD.  The default state is the reference's: the NEU flag (boundary value 0)
 * with DirGaussSeidel.  verbosity (default true, as the reference) prints the reference's
 * per-iteration record to std::clog (imp:1655-1658, 1786-1793, 1835-1844).
 */
#ifndef TK_RSFE_SPLITBREGMAN_HXX
#define TK_RSFE_SPLITBREGMAN_HXX

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "convolution.hxx"
#include "getpot.hxx"
#include "poisson.hxx"

namespace segm {

template <typename T>
class segmentation {
public:
    virtual ~segmentation() {}
    virtual void apply(im3d::image3d<T> const& myim, im3d::image3d<T> const& init) = 0;
    virtual void apply(im3d::image3d<T> const& myim) = 0;
    im3d::image3d<T> getoutput() const { return phi; }
    virtual void set_alpha(T const& alpha) { this->alpha = alpha; }
    T get_alpha() const { return alpha; }

protected:
    im3d::image3d<T> phi;
    T alpha = 0.5;
};

template <typename T>
class rsfe_splitbregman : public segmentation<T> {
public:
    enum bc_type { NEU = 1, DIR };
    enum ed_type { GRAD = 1, GRAD_THRESHOLD, LBP };

    // constructor defaults (rsfe_splitbregman_imp.hxx:25-41)
    rsfe_splitbregman() : cv_(conv::acyclic_fftw_convolution<T>), onestep_(lapl::DirGaussSeidel<T>)
    {
        tk_rsfe_default_params(&p_);
        this->alpha = (T)0.5;
    }

    // ---- setters (rsfe_splitbregman_imp.hxx:990-1330) ------------------------------------
    void set_verbosity(bool const v) { verbosity_ = v; }
    void set_maxiter(unsigned const& maxiter) { p_.maxiter = (int)maxiter; }
    void set_tol(T const& tol)
    {
        if (tol >= 0) p_.tol = (double)tol;
        else tk::warning("set_tol: tol must be non negative");
    }
    void set_lamda(T const& v) { positive(v, p_.lamda, "set_lamda: lamda must be greater than zero"); }
    void set_lamda1(T const& v) { positive(v, p_.lamda1, "set_lamda1: lamda1 must be greater than zero"); }
    void set_lamda2(T const& v) { positive(v, p_.lamda2, "set_lamda2: lamda2 must be greater than zero"); }
    void set_sigma(T const& v) { positive(v, p_.sigma, "set_sigma: sigma must be greater than zero"); }
    void set_dt(T const& v) { positive(v, p_.dt, "set_dt: dt must be greater than zero"); }
    void set_epsilon(T const& v) { positive(v, p_.epsilon, "set_epsilon: epsilon must be greater than zero"); }
    void set_nu(T const& v)
    {
        if (v >= 0) p_.nu = (double)v;
        else tk::warning("set_nu: nu must be non negative");
    }
    void set_beta(T const& v)
    {
        if (v >= 0) p_.beta = (double)v;
        else tk::warning("set_beta: beta must be non negative");
    }
    // a0 / b0 keep b0 > a0 by moving the other bound (imp:1220-1280)
    void set_a0(T const& a0)
    {
        p_.a0 = (double)a0;
        this->alpha = (T)((p_.a0 + p_.b0) / 2.);
        if (p_.b0 < p_.a0) {
            p_.b0 = p_.a0 + 1;
            this->alpha = (T)((p_.a0 + p_.b0) / 2.);
            tk::warning("set_a0: b0 has to be greater than a0, hence its value is automatically "
                        "modified");
        }
    }
    void set_b0(T const& b0)
    {
        p_.b0 = (double)b0;
        this->alpha = (T)((p_.a0 + p_.b0) / 2.);
        if (p_.b0 < p_.a0) {
            p_.a0 = p_.b0 - 1;
            this->alpha = (T)((p_.a0 + p_.b0) / 2.);
            tk::warning("set_b0: a0 has to be lower than b0, hence its value is automatically "
                        "modified");
        }
    }
    void set_ls_tol(T const& tol)
    {
        if (tol >= 0) p_.ls_tol = (double)tol;
        else tk::warning("set_ls_tol: ls_tol must be non negative");
    }
    void set_ls_steps(unsigned const steps) { p_.ls_steps = (int)steps; }
    // automatic connected component (imp:903-911, 968-1030, 1380-1417) and alpha (1422-1440)
    void set_auto_extract_conn_comp(bool const auto_cc) { p_.auto_extract_conn_comp = auto_cc ? 1 : 0; }
    void set_auto_extract_conn_comp_freq(unsigned const& cc_freq)
    {
        p_.auto_extract_conn_comp_freq = (int)cc_freq;
    }
    void set_cc_init_pixel_x(double const& v) { pixel(0, v, "set_cc_init_pixel_x"); }
    void set_cc_init_pixel_y(double const& v) { pixel(1, v, "set_cc_init_pixel_y"); }
    void set_cc_init_pixel_z(double const& v) { pixel(2, v, "set_cc_init_pixel_z"); }
    void set_cc_binary_output(bool const cc_bo) { p_.cc_binary_output = cc_bo ? 1 : 0; }
    void set_cc_init_variables(bool const cc_iv) { p_.cc_init_variables = cc_iv ? 1 : 0; }
    void set_alpha(T const& alpha) override
    {
        if (alpha > (T)p_.a0 && alpha < (T)p_.b0) this->alpha = alpha;
        else tk::warning("set_alpha: alpha must be greater than a0 and less than b0");
    }
    // dumps of phi as MetaImage files (imp:916-941, 1939-1978)
    void set_dumpfrequency(unsigned const& dump) { p_.dumpfrequency = (int)dump; }
    void set_outputname(std::string const& name)
    {
        if (name.size() < sizeof p_.outputname) {
            std::memset(p_.outputname, 0, sizeof p_.outputname);
            std::memcpy(p_.outputname, name.data(), name.size());
        } else {
            tk::warning("set_outputname: at most 255 characters; kept");
        }
    }
    void set_gaussian_pixel_approach(bool const gpa) { p_.gaussian_pixel_approach = gpa ? 1 : 0; }
    void set_edge_detector(ed_type const ed)
    {
        if (ed == GRAD) return;
        if (ed == GRAD_THRESHOLD || ed == LBP)
            tk::warning("set_edge_detector: the interactive GRAD_THRESHOLD and LBP detectors are "
                        "not built; GRAD is kept");
        else tk::warning("set_edge_detector: specify a proper ed_type condition");
    }
    // edge detector on the image blurred with this sigma (imp:1135-1153, 214-227)
    void set_edge_detector_sigma(T const& eds)
    {
        if (eds >= 0) p_.edge_detector_sigma = (double)eds;
        else tk::warning("set_edge_detector_sigma: edge_detector_sigma must be greater than zero");
    }
    // parameter file (imp:807-813) read by apply (imp:1559-1571), GetPot section
    // splitbregman/init/, and every showfrequency iterations (not at maxiter) section
    // splitbregman/onthego/ when onthego is true (update_param_onthego, imp:545-578, called at
    // 1816-1829); nothing is displayed, and without a parameter file there is no interactive
    // prompt
    void set_getpotfile(std::string const& name) { getpotfile_ = name; }
    void set_onthego(bool const onthego) { onthego_ = onthego; }
    void set_showfrequency(unsigned const& sf) { showfrequency_ = sf; }

    // set_param_from_getpot (imp:582-796): every name present in section
    // splitbregman/<section> of the parameter file goes through its setter, with the
    // reference's "absent" sentinels
    void set_param_from_getpot(std::string const& section)
    {
        tk::getpot f(getpotfile_.c_str(), "#", "\n");
        f.set_prefix(("splitbregman/" + section).c_str());
        int it;
        double t;
        std::string st;
        if ((it = f("verbosity", -1)) != -1) set_verbosity((bool)it);
        if ((it = f("maxiter", -1)) > 0) set_maxiter((unsigned)it);
        if ((it = f("showfrequency", -1)) >= 0) set_showfrequency((unsigned)it);
        if ((it = f("dumpfrequency", -1)) >= 0) set_dumpfrequency((unsigned)it);
        if ((st = f("outputname", "")).size() != 0) set_outputname(st);
        if ((it = f("save_current", -1)) != -1) save_current_ = it != 0;
        if ((it = f("end_now", -1)) != -1) end_now_ = it != 0;
        if ((t = f("sigma", -1.)) != -1.) set_sigma((T)t);
        if ((t = f("lamda", -1.)) != -1.) set_lamda((T)t);
        if ((t = f("lamda1", -1.)) != -1.) set_lamda1((T)t);
        if ((t = f("lamda2", -1.)) != -1.) set_lamda2((T)t);
        if ((t = f("nu", -1.)) != -1.) set_nu((T)t);
        if ((t = f("beta", -1.)) != -1.) set_beta((T)t);
        if ((t = f("dt", -1.)) != -1.) set_dt((T)t);
        if ((it = f("ls_steps", -1)) != -1) set_ls_steps((unsigned)it);
        if ((t = f("epsilon", -1.)) != -1.) set_epsilon((T)t);
        if ((t = f("a0", 2.1e-6)) != 2.1e-6) set_a0((T)t);
        if ((t = f("b0", -2.1e-6)) != -2.1e-6) set_b0((T)t);
        if ((it = f("auto_extract_conn_comp", -1)) != -1) set_auto_extract_conn_comp(it != 0);
        if ((it = f("cc_binary_output", -1)) != -1) set_cc_binary_output(it != 0);
        if ((it = f("cc_init_variables", -1)) != -1) set_cc_init_variables(it != 0);
        if ((t = f("cc_init_pixel_x", -1.)) != -1.) set_cc_init_pixel_x(t);
        if ((t = f("cc_init_pixel_y", -1.)) != -1.) set_cc_init_pixel_y(t);
        if ((t = f("cc_init_pixel_z", -1.)) != -1.) set_cc_init_pixel_z(t);
        if ((it = f("auto_extract_conn_comp_freq", -1)) >= 0)
            set_auto_extract_conn_comp_freq((unsigned)it);
        if ((it = f("bc", -1)) != -1) set_bc(static_cast<bc_type>(it));
        if ((t = f("ls_tol", -1.)) != -1.) set_ls_tol((T)t);
        if ((it = f("edge_detector", -1)) != -1) set_edge_detector(static_cast<ed_type>(it));
        if ((t = f("edge_detector_sigma", -1.)) != -1.) set_edge_detector_sigma((T)t);
        if ((it = f("gaussian_pixel_approach", -1)) != -1) set_gaussian_pixel_approach(it != 0);
        if ((t = f("tol", -1.)) != -1.) set_tol((T)t);
        if ((t = f("alpha", 1.e-9)) != 1.e-9) set_alpha((T)t);
    }
    // bc flag and sweep (imp:1294-1322): NEU -> NeuGaussSeidel with derivative 0, DIR ->
    // DirGaussSeidel with boundary value a0 (imp:1735-1743)
    void set_bc(bc_type const bc)
    {
        if (bc == DIR) {
            p_.bc_dir = 1;
            onestep_.changef(lapl::DirGaussSeidel<T>);
        } else if (bc == NEU) {
            p_.bc_dir = 0;
            onestep_.changef(lapl::NeuGaussSeidel<T>);
        } else {
            tk::warning("set_bc: specify a proper bc_type condition");
        }
    }
    // any convolution function (imp:836-842); a null function prints a WARNING and keeps
    void set_cv(conv::filtering<T> const cv)
    {
        if (cv.getf()) cv_ = cv;
        else tk::warning("set_cv: null function; kept");
    }
    // any one-sweep Poisson function (imp:845-852)
    void set_onestep_poisson(lapl::unsteady_poisson_functor<T> const osl)
    {
        if (osl.getf()) onestep_ = osl;
        else tk::warning("set_onestep_poisson: null function; kept");
    }
    // true when apply runs the fused device loop, false for the stage-by-stage path that
    // calls the installed functors
    bool fused() const
    {
        const bool cv = cv_.getf() == &conv::cyclic_fftw_convolution<T> ||
                        cv_.getf() == &conv::acyclic_fftw_convolution<T>;
        const bool gs = onestep_.getf() == &lapl::DirGaussSeidel<T> ||
                        onestep_.getf() == &lapl::DirRedBlackGaussSeidel<T> ||
                        onestep_.getf() == &lapl::NeuGaussSeidel<T>;
        return cv && gs;
    }
    // cap on the sweeps of one Poisson solve when ls_tol > 0 (0: none)
    void set_max_sweeps(int const m) { p_.max_sweeps = m; }

    // ---- apply (rsfe_splitbregman_imp.hxx:1547-1995) --------------------------------------
    void apply(im3d::image3d<T> const& myim, im3d::image3d<T> const& init) override
    {
        run(myim, init.handle());
    }
    // initial contour: the cube (initialize_phi_with_cube, imp:48-95)
    void apply(im3d::image3d<T> const& myim) override { run(myim, nullptr); }

    // per outer iteration: Gauss-Seidel sweeps, tollerance (imp:1832)
    std::vector<int> const& get_sweeps() const { return sweeps_; }
    std::vector<double> const& get_tollerances() const { return tols_; }
    unsigned get_iterations() const { return (unsigned)sweeps_.size(); }
    tk_rsfe_params const& params() const { return p_; }

private:
    void pixel(int a, double const& v, const char* name)
    {
        if (v >= 0) p_.cc_init_pixel[a] = v;
        else tk::warning(std::string(name) + ": coordinates must be positive");
    }
    void positive(T const& v, double& slot, const char* msg)
    {
        if (v > 0) slot = (double)v;
        else tk::warning(msg);
    }

    // the device run's parameters from the current settings
    tk_rsfe_params device_params() const
    {
        tk_rsfe_params p = p_;
        p.gs_order = onestep_.getf() == &lapl::DirRedBlackGaussSeidel<T> ? TK_GS_REDBLACK
                                                                           : TK_GS_LEX;
        p.acyclic = cv_.getf() == &conv::acyclic_fftw_convolution<T> ? 1 : 0;
        p.neumann = onestep_.getf() == &lapl::NeuGaussSeidel<T> ? 1 : 0;
        p.verbosity = verbosity_ ? 1 : 0;
        // extract_connected_component's threshold (imp:481), in T as the reference
        p.cc_threshold = (double)((this->alpha - (T)p_.a0) / ((T)p_.b0 - (T)p_.a0));
        return p;
    }

    void run(im3d::image3d<T> const& myim, tk_vol init)
    {
        // the parameter file's init section first (imp:1559-1571)
        save_current_ = end_now_ = false;
        if (getpotfile_ != "") set_param_from_getpot("init/");
        else onthego_ = false;
        tk_rsfe_params p = device_params();
        tk_rsfe s = nullptr;
        sweeps_.clear();
        tols_.clear();
        if (!fused()) {
            run_generic(myim, init, p);
            return;
        }
        tk::check(tk_rsfe_create(myim.ctx(), myim.handle(), init, &p, &s), "apply");
        // with the on-the-go re-read, chunks of showfrequency iterations; after each one
        // (t % showfrequency == 0, t != maxiter) section splitbregman/onthego/ of the file
        const bool onthego = onthego_ && showfrequency_ > 0;
        int r = TK_OK, t = 0;
        for (;;) {
            const int left = p.maxiter > t ? p.maxiter - t : 0;
            const int want = onthego ? std::min((int)showfrequency_, left) : left;
            std::vector<tk_rsfe_iter_stats> st((size_t)(want > 0 ? want : 1));
            int done = 0;
            r = tk_rsfe_iterate(s, want, st.data(), &done);
            for (int q = 0; q < done; ++q) {
                sweeps_.push_back(st[q].sweeps);
                tols_.push_back(st[q].tollerance);
                if (verbosity_) log_iteration(t + q + 1, st[q].sweeps, st[q].ls_err,
                                              st[q].tol_inf, st[q].tollerance);
            }
            t += done;
            if (r != TK_OK || !onthego || done < want || want == 0 || t == p.maxiter) break;
            if (verbosity_)
                std::clog << "-- Updating parameters of algorithm from file " << getpotfile_
                          << " (section splitbregman/onthego/)" << std::endl;
            set_param_from_getpot("onthego/");
            tk_rsfe_params q = device_params();
            if (q.gs_order != p.gs_order || q.acyclic != p.acyclic || q.neumann != p.neumann) {
                tk::warning("update_param_onthego: the Poisson solver and cv of a running apply "
                            "are kept");
                q.gs_order = p.gs_order;
                q.acyclic = p.acyclic;
                q.neumann = p.neumann;
            }
            p = q;
            const bool end = end_now_;
            r = tk_rsfe_onthego(s, &p, save_current_ ? 1 : 0, end ? 1 : 0);
            save_current_ = false;
            if (r != TK_OK) break;
            if (end) {
                // iterations end here; the final dump is written (imp:1659, 1968-1978)
                end_now_ = false;
                break;
            }
        }

        if (r == TK_OK) {
            this->phi = im3d::image3d<T>(myim.getdimx(), myim.getdimy(), myim.getdimz(),
                                         myim.gethx(), myim.gethy(), myim.gethz());
            r = tk_vol_copy(this->phi.handle(), tk_rsfe_phi(s));
        }
        tk_rsfe_destroy(s);
        tk::check(r, "apply");
    }

    // the reference's verbose record of one outer iteration (imp:1655-1658, 1786-1793,
    // 1835-1844)
    static void log_iteration(int t, int sweeps, double ls_err, double tol_inf, double tol)
    {
        std::clog << "-- Iteration " << t << std::endl;
        std::clog << "\tlinear system error at " << sweeps << "th (last) step: " << (T)ls_err
                  << std::endl;
        std::clog << "\tnorm_inf: " << (T)tol_inf << "\tnorm_L2: " << (T)tol << std::endl;
    }

    // apply stage by stage (rsfe_splitbregman_imp.hxx:1641-1936) on device image3d objects,
    // with the installed cv and onestep_poisson functors called where the reference calls
    // them and every other step in the reference's order
    void run_generic(im3d::image3d<T> const& myim, tk_vol init, tk_rsfe_params const& p)
    {
        if (p.auto_extract_conn_comp || p.dumpfrequency)
            tk::warning("apply: the automatic connected component and the dumps run with the "
                        "built-in cv and onestep_poisson only; ignored");
        tk_ctx ctx = myim.ctx();
        const uint X = myim.getdimx(), Y = myim.getdimy(), Z = myim.getdimz();
        const double hx = myim.gethx(), hy = myim.gethy(), hz = myim.gethz();
        auto field = [&]() { return im3d::image3d<T>(X, Y, Z, hx, hy, hz); };
        // phi0: initialize_phi_with_init / _with_cube (imp:48-170), the fused path's own
        im3d::image3d<T> phi = field();
        {
            tk_rsfe_params q = p;
            q.verbosity = 0;
            tk_rsfe s0 = nullptr;
            tk::check(tk_rsfe_create(ctx, myim.handle(), init, &q, &s0), "apply");
            const int r = tk_vol_copy(phi.handle(), tk_rsfe_phi(s0));
            tk_rsfe_destroy(s0);
            tk::check(r, "apply");
        }
        // filter, kones = cv(1), kmyim = cv(myim) (imp:1641-1651)
        conv::filtering<T>::build_gaussian_filter(phi, (T)p.sigma, p.radius,
                                                  p.gaussian_pixel_approach != 0);
        im3d::image3d<T> kones, kmyim, aux = field();
        aux = (T)1;
        cv_(kones, aux);
        cv_(kmyim, myim);
        // gforshrink = edge_detect(myim) (imp:196-229, 1655); with edge_detector_sigma the
        // image is first convolved by cv with that sigma, then the filter is rebuilt
        im3d::image3d<T> g = field();
        if (p.edge_detector_sigma != 0) {
            conv::filtering<T>::build_gaussian_filter(phi, (T)p.edge_detector_sigma, 3,
                                                      p.gaussian_pixel_approach != 0);
            im3d::image3d<T> blurred;
            cv_(blurred, myim);
            tk::check(tk_edge_detect(ctx, blurred.handle(), p.beta, p.lamda, p.nu, g.handle()),
                      "edge_detect");
            conv::filtering<T>::build_gaussian_filter(phi, (T)p.sigma, p.radius,
                                                      p.gaussian_pixel_approach != 0);
        } else {
            tk::check(tk_edge_detect(ctx, myim.handle(), p.beta, p.lamda, p.nu, g.handle()),
                      "edge_detect");
        }
        std::vector<im3d::image3d<T> > b(3, field()), d;
        im3d::image3d<T> divdb = field();
        const double chosen_bc = p.bc_dir ? p.a0 : 0.;  // imp:1735-1743
        double tollerance = 1000;
        for (int t = 1; tollerance > p.tol && t <= p.maxiter; ++t) {
            im3d::image3d<T> phiold(phi);
            // region fit and forcing (imp:1690-1759)
            im3d::image3d<T> H = field(), num, den, kr1, kr2;
            tk::check(tk_heaviside(ctx, phi.handle(), p.a0, p.b0, p.epsilon, H.handle()),
                      "set_heaviside");
            cv_(num, H * myim);
            cv_(den, H);
            im3d::image3d<T> f1 = num / den;
            num -= kmyim;
            den -= kones;
            num /= den;
            im3d::image3d<T> f2 = num;
            cv_(kr2, f1 * (T)p.lamda1 - f2 * (T)p.lamda2);
            f1 *= f1;
            f1 *= (T)p.lamda1;
            f2 *= f2;
            f2 *= (T)p.lamda2;
            f1 -= f2;
            cv_(kr1, f1);
            kr2 *= myim;
            kr2 *= 2;
            aux.setdim(X, Y, Z, (T)((T)p.lamda1 - (T)p.lamda2));
            aux *= myim;
            aux *= myim;
            aux *= kones;
            kr2 -= aux;
            kr2 -= kr1;
            kr2 /= (T)p.lamda;
            kr2 -= divdb;
            // Poisson loop (imp:1762-1797); max_sweeps > 0 caps it
            T ls_err = p.ls_tol != 0 ? (T)1000 : (T)-1;
            int s;
            for (s = 0; s < p.ls_steps || ls_err > (T)p.ls_tol; ++s) {
                if (p.max_sweeps > 0 && s >= p.max_sweeps) break;
                if ((verbosity_ && s == p.ls_steps - 1) || p.ls_tol != 0) aux = phi;
                onestep_(phi, kr2, (T)p.dt, (T)chosen_bc);
                if (p.ls_tol != 0 && s >= p.ls_steps - 1)
                    ls_err = (phi - aux).normL2() / aux.normL2();
            }
            if (verbosity_ && p.ls_tol == 0) ls_err = (phi - aux).normL2() / aux.normL2();
            // cut_phi (imp:320-339), tollerance (imp:1832)
            tk::check(tk_cut(ctx, phi.handle(), p.a0, p.b0), "cut_phi");
            double dn = 0, on = 0;
            tk::check(tk_norm_l2_diff(ctx, phi.handle(), phiold.handle(), &dn), "tollerance");
            tk::check(tk_norm_l2(ctx, phiold.handle(), &on), "tollerance");
            tollerance = dn / on;
            sweeps_.push_back(s);
            tols_.push_back(tollerance);
            if (verbosity_)
                log_iteration(t, s, (double)ls_err,
                              (double)((phi - phiold).norminf() / phiold.norminf()), tollerance);
            // Split-Bregman update (imp:1868-1936)
            phi.grad(d);
            for (int i = 0; i < 3; ++i) b[i] += d[i];
            for (int i = 0; i < 3; ++i)
                tk::check(tk_shrink(ctx, b[i].handle(), g.handle(), d[i].handle()), "shrink");
            for (int i = 0; i < 3; ++i) {
                b[i] -= d[i];
                d[i] -= b[i];
            }
            im3d::div(divdb, d);
        }
        this->phi = phi;
    }

    tk_rsfe_params p_;
    bool verbosity_ = true;
    std::string getpotfile_ = "data";  // the constructor's default (imp:27)
    bool onthego_ = true;
    unsigned showfrequency_ = 0;
    bool save_current_ = false, end_now_ = false;  // set by section splitbregman/onthego/
    conv::filtering<T> cv_;
    lapl::unsteady_poisson_functor<T> onestep_;
    std::vector<int> sweeps_;
    std::vector<double> tols_;
};

}  // namespace segm

#endif
