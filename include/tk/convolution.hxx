/*
 * tk/convolution.hxx -- conv::filtering<T>, conv::cyclic_fftw_convolution and
 * conv::acyclic_fftw_convolution over the device filters of libtk.
 *
 * Mirrors conv::filtering<T> (header/convolution.hxx:65-315): a functor holding a plain
 * function pointer void (*f)(image3d<T>& result, image3d<T> const& factor) (79), its
 * constructor (89-90), operator() (104-107) and changef (119-122), with the filter kept in
 * class-wide static state built by the static build_gaussian_filter (226-229) and
 * build_freq_filters (299).  cyclic_fftw_convolution (391; convolution_imp.hxx:657-757) is
 * the circular convolution of factor with that filter: it resizes result to the factor's
 * dims and copies its spacing (convolution_imp.hxx:678-686).  acyclic_fftw_convolution
 * (368; convolution_imp.hxx:535-653) is the convolution with zero values outside the
 * volume (the reference pads to dims + Nh - 1 and keeps the first dims voxels).
 *
 * The static state is per thread here (the reference's is per process and not re-entrant):
 * the Gaussian parameters of the last build_gaussian_filter and a device filter built for
 * the dims of the last build_freq_filters.  direct_convolution (411; convolution_imp.hxx:
 * 761-812) uses the same filter.  The reference's behaviour is undefined when the
 * factor's dims differ from the ones the filter was built for (convolution.hxx:286-297);
 * this layer prints a WARNING and rebuilds the filter for the factor's dims.
 */
#ifndef TK_CONVOLUTION_HXX
#define TK_CONVOLUTION_HXX

#include "image3d.hxx"

namespace conv {

template <typename T>
class filtering {
public:
    typedef void (*function)(im3d::image3d<T>& result, im3d::image3d<T> const& factor);

    explicit filtering(function f) : f_(f) {}

    void operator()(im3d::image3d<T>& result, im3d::image3d<T> const& factor) const
    {
        f_(result, factor);
    }

    void changef(function newf) { f_ = newf; }
    function getf() const { return f_; }

    // Gaussian kernel exp(-r^2 / (2 sigma^2)) on Nh^3 taps normalised to unit L1 norm
    // (convolution_imp.hxx:22-91), then build_freq_filters(f) as the reference does
    static void build_gaussian_filter(im3d::image3d<T> const& f, T const& sigma,
                                      double const& radius = 3, bool pixel_approach = false)
    {
        if (!(sigma > 0)) {
            tk::warning("build_gaussian_filter: sigma must be greater than zero");
            return;
        }
        state& s = st();
        s.sigma = (double)sigma;
        s.radius = radius;
        s.pixel = pixel_approach;
        s.have_gaussian = true;
        build_freq_filters(f);
    }

    // spectrum of the current filter for f's dims and spacing (convolution_imp.hxx:218-531,
    // cyclic part)
    static void build_freq_filters(im3d::image3d<T> const& f)
    {
        state& s = st();
        if (!s.have_gaussian) {
            tk::warning("build_freq_filters: no time filter built (call build_gaussian_filter)");
            return;
        }
        s.release();
        tk::check(tk_filter_gaussian(f.ctx(), (int)f.getdimx(), (int)f.getdimy(),
                                     (int)f.getdimz(), f.gethx(), f.gethy(), f.gethz(), s.sigma,
                                     s.radius, s.pixel ? 1 : 0, tk::dtype_of<T>::value, &s.flt),
                  "build_freq_filters");
        s.dims[0] = f.getdimx();
        s.dims[1] = f.getdimy();
        s.dims[2] = f.getdimz();
        s.h[0] = f.gethx();
        s.h[1] = f.gethy();
        s.h[2] = f.gethz();
        s.ctx = f.ctx();
    }

    // the device filter of the static state, rebuilt for factor's dims if they differ
    static tk_filter filter_for(im3d::image3d<T> const& factor)
    {
        state& s = st();
        if (!s.flt || s.ctx != factor.ctx() || s.dims[0] != factor.getdimx() ||
            s.dims[1] != factor.getdimy() || s.dims[2] != factor.getdimz() ||
            s.h[0] != factor.gethx() || s.h[1] != factor.gethy() || s.h[2] != factor.gethz()) {
            if (s.flt)
                tk::warning("fftw_convolution: factor dims differ from the filter's; "
                            "rebuilding the frequency filter");
            build_freq_filters(factor);
        }
        return s.flt;
    }

private:
    struct state {
        tk_filter flt = nullptr;
        tk_ctx ctx = nullptr;
        unsigned dims[3] = {0, 0, 0};
        double h[3] = {0, 0, 0};
        double sigma = 0, radius = 3;
        bool pixel = false, have_gaussian = false;
        void release()
        {
            if (flt) tk_filter_free(flt);
            flt = nullptr;
        }
        ~state() { release(); }
    };
    static state& st()
    {
        thread_local state s;
        return s;
    }
    function f_;
};

// result = factor (*) filter, circular boundaries (convolution_imp.hxx:657-757)
template <typename T>
void cyclic_fftw_convolution(im3d::image3d<T>& result, im3d::image3d<T> const& factor)
{
    tk_filter flt = filtering<T>::filter_for(factor);
    if (!flt) return;
    result.shape_like(factor);
    tk::check(tk_conv_cyclic(factor.ctx(), flt, result.handle(), factor.handle()),
              "cyclic_fftw_convolution");
}

// result = factor (*) filter, zero outside the volume (convolution_imp.hxx:535-653)
template <typename T>
void acyclic_fftw_convolution(im3d::image3d<T>& result, im3d::image3d<T> const& factor)
{
    tk_filter flt = filtering<T>::filter_for(factor);
    if (!flt) return;
    result.shape_like(factor);
    tk::check(tk_conv_acyclic(factor.ctx(), flt, result.handle(), factor.handle()),
              "acyclic_fftw_convolution");
}

// direct_convolution (convolution.hxx:411; convolution_imp.hxx:761-812): result(i,j,k) =
// sum over the Nh^3 taps of hx hy hz * time_filter(I,J,K) * factor, zero outside the volume.
// The Gaussian time filter is a product of per-axis factors, so the triple sum equals three
// 1-D passes with zero halos, which is what runs (the acyclic passes of the device filter);
// the result equals the reference's up to rounding, not in its summation order.
template <typename T>
void direct_convolution(im3d::image3d<T>& result, im3d::image3d<T> const& factor)
{
    tk_filter flt = filtering<T>::filter_for(factor);
    if (!flt) return;
    result.shape_like(factor);
    tk::check(tk_conv_acyclic(factor.ctx(), flt, result.handle(), factor.handle()),
              "direct_convolution");
}

}  // namespace conv

#endif
