/*
 * tk/image3d.hxx -- im3d::image3d<T> with its voxels resident in device memory.
 *
 * Mirrors im3d::image3d<T> (header/image3d.hxx:50-737) for what the RSFE Split-Bregman path
 * uses: the constructors (image3d.hxx:90-111), dims and spacing accessors (122-238),
 * element read and write (252, 262), operator=(scalar) and operator=(image3d) (271-280),
 * the compound operators += -= *= /= with a scalar or an image and the binary + - * /
 * (283-402), scalarprod / scalarprod_L2 (420-430), grad (442), vector_abs (448), div (454),
 * norm1 / normL1 / norm2 / normL2 / norminf / min / max (480-513), im_to_black_and_white
 * (621), change_range_of_intensity (635) and the seeded connected_component (679).  Value
 * semantics as the reference (copies are deep, image3d_imp.hxx:44-57).  Voxel order is the
 * reference's: (i*dimy + j)*dimz + k (image3d_imp.hxx:141-152), so upload() and download()
 * take and give the reference's raw vector unchanged.  T is double, float or unsigned char
 * (masks).
 */
#ifndef TK_IMAGE3D_HXX
#define TK_IMAGE3D_HXX

#include <cmath>
#include <cstdint>
#include <iostream>
#include <type_traits>
#include <utility>
#include <vector>

#include "context.hxx"

namespace im3d {

typedef unsigned int uint;

// the reference's image3d messages go to std::cout (image3d_imp.hxx:220-223, 400-426)
inline void image3d_warning(const char* msg) { std::cout << "WARNING::" << msg << std::endl; }

template <typename T>
class image3d;

// a writable voxel, what operator() returns for a non-const image (image3d.hxx:262: T&):
// reads and writes go to device memory one voxel at a time
template <typename T>
class voxel_ref {
public:
    voxel_ref(tk_vol v, uint i, uint j, uint k) : v_(v), i_(i), j_(j), k_(k) {}
    operator T() const
    {
        double x = 0;
        if (v_) tk::check(tk_vol_get(v_, (int)i_, (int)j_, (int)k_, &x), "image3d element");
        return (T)x;
    }
    voxel_ref& operator=(T const& x)
    {
        if (v_) tk::check(tk_vol_set(v_, (int)i_, (int)j_, (int)k_, (double)x), "image3d element");
        return *this;
    }
    voxel_ref& operator=(voxel_ref const& o) { return *this = (T)o; }
    voxel_ref& operator+=(T const& x) { return *this = (T)(T(*this) + x); }
    voxel_ref& operator-=(T const& x) { return *this = (T)(T(*this) - x); }
    voxel_ref& operator*=(T const& x) { return *this = (T)(T(*this) * x); }
    voxel_ref& operator/=(T const& x) { return *this = (T)(T(*this) / x); }

private:
    tk_vol v_;
    uint i_, j_, k_;
};

template <typename T>
class image3d {
public:
    image3d() : ctx_(tk::context::current_ptr()) {}

    image3d(uint const& x, uint const& y, uint const& z, double const& hx = 1.,
            double const& hy = 1., double const& hz = 1.)
        : ctx_(tk::context::current_ptr()), hx_(hx), hy_(hy), hz_(hz)
    {
        alloc(x, y, z);  // zero-filled, as the reference's std::vector
    }

    image3d(image3d const& o) : ctx_(o.ctx_), hx_(o.hx_), hy_(o.hy_), hz_(o.hz_)
    {
        alloc(o.dimx_, o.dimy_, o.dimz_);
        if (v_) tk::check(tk_vol_copy(v_, o.v_), "image3d copy");
    }

    image3d(image3d&& o) noexcept : ctx_(o.ctx_) { swap(o); }

    image3d& operator=(image3d const& o)
    {
        if (this == &o) return *this;
        if (o.dimx_ != dimx_ || o.dimy_ != dimy_ || o.dimz_ != dimz_ || ctx_ != o.ctx_) {
            release();
            ctx_ = o.ctx_;
            alloc(o.dimx_, o.dimy_, o.dimz_);
        }
        seth(o.hx_, o.hy_, o.hz_);
        if (v_) tk::check(tk_vol_copy(v_, o.v_), "image3d copy");
        return *this;
    }

    image3d& operator=(image3d&& o) noexcept
    {
        swap(o);
        return *this;
    }

    // operator=(S const&): every voxel set to the value (image3d_imp.hxx:158-167)
    image3d& operator=(T const& value)
    {
        if (v_) tk::check(tk_vol_fill(v_, (double)value), "image3d fill");
        return *this;
    }

    // conversion between element types (operator=(image3d<S>), image3d_imp.hxx:170-184)
    template <typename S>
    image3d& assign(image3d<S> const& o)
    {
        if (o.getdimx() != dimx_ || o.getdimy() != dimy_ || o.getdimz() != dimz_) {
            release();
            alloc(o.getdimx(), o.getdimy(), o.getdimz());
        }
        seth(o.gethx(), o.gethy(), o.gethz());
        if (v_) tk::check(tk_vol_copy(v_, o.handle()), "image3d convert");
        return *this;
    }

    ~image3d() { release(); }

    // ---- dims and spacing -------------------------------------------------------
    uint getdimx() const { return dimx_; }
    uint getdimy() const { return dimy_; }
    uint getdimz() const { return dimz_; }
    double gethx() const { return hx_; }
    double gethy() const { return hy_; }
    double gethz() const { return hz_; }

    // setdim frees and reallocates filled with value (image3d_imp.hxx:93-103)
    void setdim(uint const& x, uint const& y, uint const& z, T const& value = 0)
    {
        release();
        alloc(x, y, z);
        if (v_ && value != T(0)) *this = value;
    }
    void seth(double const& hx, double const& hy, double const& hz)
    {
        hx_ = hx;
        hy_ = hy;
        hz_ = hz;
        if (v_) tk::check(tk_vol_set_spacing(v_, hx, hy, hz), "image3d seth");
    }
    void sethx(double const& hx) { seth(hx, hy_, hz_); }
    void sethy(double const& hy) { seth(hx_, hy, hz_); }
    void sethz(double const& hz) { seth(hx_, hy_, hz); }

    // ---- host transfer (the reference's rawimage) ---------------------------------
    void upload(std::vector<T> const& raw)
    {
        if (raw.size() != size()) {
            tk::warning("image3d::upload: size of the data does not match the dims");
            return;
        }
        if (v_) tk::check(tk_vol_upload(v_, raw.data(), tk::dtype_of<T>::value), "upload");
    }
    std::vector<T> download() const
    {
        std::vector<T> raw(size());
        if (v_) tk::check(tk_vol_download(v_, raw.data(), tk::dtype_of<T>::value), "download");
        return raw;
    }

    // element read (operator() const, image3d.hxx:252): one device-to-host copy
    T operator()(uint const& i, uint const& j, uint const& k) const
    {
        T x = T(0);
        if (!v_ || i >= dimx_ || j >= dimy_ || k >= dimz_) return x;
        tk_vol one = nullptr;
        const long off = ((long)i * dimy_ + j) * dimz_ + k;
        tk::check(tk_vol_wrap(ctx_->handle(), (T*)tk_vol_device_ptr(v_) + off, 1, 1, 1, 1, 1, 1,
                              tk::dtype_of<T>::value, &one),
                  "image3d element");
        int st = tk_vol_download(one, &x, tk::dtype_of<T>::value);
        tk_vol_free(one);
        tk::check(st, "image3d element");
        return x;
    }

    // element write (operator(), image3d.hxx:262): one host-to-device copy per assignment
    voxel_ref<T> operator()(uint const& i, uint const& j, uint const& k)
    {
        return voxel_ref<T>(i < dimx_ && j < dimy_ && k < dimz_ ? v_ : nullptr, i, j, k);
    }

    // ---- compound operators (image3d_imp.hxx:187-345) -----------------------------
    // with a scalar: static_cast<T>(s), then every voxel in T arithmetic
    template <typename S, typename std::enable_if<std::is_arithmetic<S>::value, int>::type = 0>
    image3d& operator+=(S const& s) { return op_scalar(TK_OP_ADD, s); }
    template <typename S, typename std::enable_if<std::is_arithmetic<S>::value, int>::type = 0>
    image3d& operator-=(S const& s) { return op_scalar(TK_OP_SUB, s); }
    template <typename S, typename std::enable_if<std::is_arithmetic<S>::value, int>::type = 0>
    image3d& operator*=(S const& s) { return op_scalar(TK_OP_MUL, s); }
    template <typename S, typename std::enable_if<std::is_arithmetic<S>::value, int>::type = 0>
    image3d& operator/=(S const& s) { return op_scalar(TK_OP_DIV, s); }
    // with an image of the same dims, voxel by voxel (WARNING and nothing done otherwise)
    template <typename S>
    image3d& operator+=(image3d<S> const& o) { return op_image(TK_OP_ADD, o, "image3d::operator+= : dimensions must agree"); }
    template <typename S>
    image3d& operator-=(image3d<S> const& o) { return op_image(TK_OP_SUB, o, "image3d::operator-= : dimensions must agree"); }
    template <typename S>
    image3d& operator*=(image3d<S> const& o) { return op_image(TK_OP_MUL, o, "image3d::operator*= : dimensions must agree"); }
    template <typename S>
    image3d& operator/=(image3d<S> const& o) { return op_image(TK_OP_DIV, o, "image3d::operator/= : dimensions must agree"); }

    // ---- norms (image3d_imp.hxx:1387-1452), FP64 accumulation --------------------------
    T const norm1() const
    {
        double r = 0;
        tk::check(tk_norm1(ctx_->handle(), v_, &r), "norm1");
        return (T)r;
    }
    T const normL1() const { return (T)((double)norm1() * hx_ * hy_ * hz_); }
    T const norm2() const
    {
        double r = 0;
        tk::check(tk_scalarprod(ctx_->handle(), v_, v_, &r), "norm2");
        return (T)std::sqrt(r);
    }
    // the reference's quirk: entries below the running maximum, the negative ones among
    // them, never count (image3d_imp.hxx:1439)
    T const norminf() const
    {
        double r = 0;
        tk::check(tk_norminf(ctx_->handle(), v_, &r), "norminf");
        return (T)r;
    }

    // ---- operators of the path ---------------------------------------------------
    // image3d::grad (image3d_imp.hxx:743-888)
    void grad(std::vector<image3d<T> >& res) const
    {
        res.resize(3);
        for (auto& r : res)
            if (r.getdimx() != dimx_ || r.getdimy() != dimy_ || r.getdimz() != dimz_) {
                r = image3d<T>(dimx_, dimy_, dimz_, hx_, hy_, hz_);
            }
        tk::check(tk_grad(ctx_->handle(), v_, res[0].v_, res[1].v_, res[2].v_), "grad");
    }

    // normL2 = sqrt(hx hy hz sum x^2) (image3d_imp.hxx:1419-1423), FP64 accumulation
    T const normL2() const
    {
        double r = 0;
        tk::check(tk_norm_l2(ctx_->handle(), v_, &r), "normL2");
        return (T)r;
    }
    T const max() const
    {
        double mn = 0, mx = 0;
        tk::check(tk_minmax(ctx_->handle(), v_, &mn, &mx), "max");
        return (T)mx;
    }
    T const min() const
    {
        double mn = 0, mx = 0;
        tk::check(tk_minmax(ctx_->handle(), v_, &mn, &mx), "min");
        return (T)mn;
    }

    // image3d_imp.hxx:1627-1635
    void change_range_of_intensity(T const& max, T const& min = 0)
    {
        tk::check(tk_change_range_of_intensity(ctx_->handle(), v_, (double)max, (double)min),
                  "change_range_of_intensity");
    }

    // image3d_imp.hxx:1591-1623; res is a 0/1 mask
    void im_to_black_and_white(image3d<unsigned char>& res, double threshold = 0.5,
                               bool negative = false) const
    {
        if (threshold > 1 || threshold < 0) {
            // image3d_imp.hxx:1597-1602
            tk::warning("im_to_black_and_white: threshold must be between 0 and 1, otherwise "
                        "default value 0.5 is applied");
            threshold = 0.5;
        }
        res.shape_like(*this);
        tk::check(tk_im_to_black_and_white(ctx_->handle(), v_, threshold, negative ? 1 : 0,
                                           res.handle()),
                  "im_to_black_and_white");
    }

    // connected_component(res, i, j, k, threshold, full_connected, binary_output)
    // (image3d_imp.hxx:1694-1757); binary output only
    void connected_component(image3d<unsigned char>& res, uint const& i, uint const& j,
                             uint const& k, double threshold = 0.5, bool full_connected = false,
                             bool binary_output = true) const
    {
        if (!binary_output) {
            tk::warning("connected_component: only binary output is built");
            return;
        }
        res.shape_like(*this);
        tk::check(tk_cc_seeded(ctx_->handle(), v_, (int)i, (int)j, (int)k, threshold,
                               full_connected ? 1 : 0, res.handle(), nullptr),
                  "connected_component");
    }

    // ---- access to the device volume ------------------------------------------------
    tk_vol handle() const { return v_; }
    tk_ctx ctx() const { return ctx_->handle(); }
    size_t size() const { return (size_t)dimx_ * dimy_ * dimz_; }

    template <typename S>
    void shape_like(image3d<S> const& o)
    {
        if (o.getdimx() != dimx_ || o.getdimy() != dimy_ || o.getdimz() != dimz_ ||
            o.gethx() != hx_ || o.gethy() != hy_ || o.gethz() != hz_) {
            release();
            hx_ = o.gethx();
            hy_ = o.gethy();
            hz_ = o.gethz();
            alloc(o.getdimx(), o.getdimy(), o.getdimz());
        }
    }

private:
    template <typename S>
    image3d& op_scalar(int op, S const& s)
    {
        if (v_) tk::check(tk_vol_op_scalar(ctx_->handle(), v_, op, (double)(T)s), "image3d operator");
        return *this;
    }
    template <typename S>
    image3d& op_image(int op, image3d<S> const& o, const char* msg)
    {
        if (o.getdimx() != dimx_ || o.getdimy() != dimy_ || o.getdimz() != dimz_) {
            image3d_warning(msg);
            return *this;
        }
        if (v_) tk::check(tk_vol_op_vol(ctx_->handle(), v_, op, o.handle()), "image3d operator");
        return *this;
    }
    void alloc(uint x, uint y, uint z)
    {
        dimx_ = x;
        dimy_ = y;
        dimz_ = z;
        v_ = nullptr;
        if ((size_t)x * y * z == 0) return;
        tk::check(tk_vol_alloc(ctx_->handle(), (int)x, (int)y, (int)z, hx_, hy_, hz_,
                               tk::dtype_of<T>::value, &v_),
                  "image3d alloc");
    }
    void release()
    {
        if (v_) tk_vol_free(v_);
        v_ = nullptr;
        dimx_ = dimy_ = dimz_ = 0;
    }
    void swap(image3d& o) noexcept
    {
        std::swap(ctx_, o.ctx_);
        std::swap(v_, o.v_);
        std::swap(dimx_, o.dimx_);
        std::swap(dimy_, o.dimy_);
        std::swap(dimz_, o.dimz_);
        std::swap(hx_, o.hx_);
        std::swap(hy_, o.hy_);
        std::swap(hz_, o.hz_);
    }

    std::shared_ptr<tk::context> ctx_;
    tk_vol v_ = nullptr;
    uint dimx_ = 0, dimy_ = 0, dimz_ = 0;
    double hx_ = 1., hy_ = 1., hz_ = 1.;
};

// binary operators (image3d_imp.hxx:349-378): a copy of the left operand, then the compound
// operator with the right one (an image or a scalar); the result has the left's type
template <typename S, typename R>
image3d<S> const operator+(image3d<S> const& a, R const& b)
{
    image3d<S> r(a);
    return r += b;
}
template <typename S, typename R>
image3d<S> const operator-(image3d<S> const& a, R const& b)
{
    image3d<S> r(a);
    return r -= b;
}
template <typename S, typename R>
image3d<S> const operator*(image3d<S> const& a, R const& b)
{
    image3d<S> r(a);
    return r *= b;
}
template <typename S, typename R>
image3d<S> const operator/(image3d<S> const& a, R const& b)
{
    image3d<S> r(a);
    return r /= b;
}

// scalarprod (image3d_imp.hxx:385-407): sum a * static_cast<S>(b), FP64 accumulation; dims
// that disagree: WARNING and 0
template <typename S, typename R>
S const scalarprod(image3d<S> const& a, image3d<R> const& b)
{
    if (a.getdimx() != b.getdimx() || a.getdimy() != b.getdimy() || a.getdimz() != b.getdimz()) {
        image3d_warning("scalarprod: dimensions must agree, 0 is returned");
        return 0;
    }
    double r = 0;
    tk::check(tk_scalarprod(a.ctx(), a.handle(), b.handle(), &r), "scalarprod");
    return (S)r;
}

// scalarprod_L2 (image3d_imp.hxx:411-426): hx hy hz scalarprod; spacings that disagree:
// WARNING and 0
template <typename S, typename R>
S const scalarprod_L2(image3d<S> const& a, image3d<R> const& b)
{
    if (a.gethx() != b.gethx() || a.gethy() != b.gethy() || a.gethz() != b.gethz()) {
        image3d_warning("scalarprod_L2: spacing must agree, 0 is returned");
        return 0;
    }
    return (S)(a.gethx() * a.gethy() * a.gethz() * (double)scalarprod(a, b));
}

// vector_abs (image3d_imp.hxx:892-943): res = |fun| voxel by voxel for 2 or 3 components
template <typename S, typename R>
void vector_abs(image3d<S>& res, std::vector<image3d<R> > const& fun)
{
    if (fun.size() != 2 && fun.size() != 3) {
        std::cout << "In image3d::vector_abs: input vector could be only 2d or 3d, "
                  << fun.size() << " is a not allowed dimension." << std::endl;
        return;
    }
    res.shape_like(fun[0]);
    for (size_t c = 1; c < fun.size(); ++c)
        if (fun[c].getdimx() != fun[0].getdimx() || fun[c].getdimy() != fun[0].getdimy() ||
            fun[c].getdimz() != fun[0].getdimz()) {
            std::cout << "In image3d::vector_abs: dimensions of image3d "
                      << "inside input vector must agree" << std::endl;
            return;
        }
    if (std::is_same<S, R>::value) {
        tk::check(tk_vector_abs(res.ctx(), fun[0].handle(), fun[1].handle(),
                                fun.size() == 3 ? fun[2].handle() : nullptr, res.handle()),
                  "vector_abs");
    } else {
        // computed in R (the components' type), then cast to S
        image3d<R> tmp(fun[0].getdimx(), fun[0].getdimy(), fun[0].getdimz(), fun[0].gethx(),
                       fun[0].gethy(), fun[0].gethz());
        tk::check(tk_vector_abs(res.ctx(), fun[0].handle(), fun[1].handle(),
                                fun.size() == 3 ? fun[2].handle() : nullptr, tmp.handle()),
                  "vector_abs");
        res.assign(tmp);
    }
}

// im3d::div (image3d_imp.hxx:947-1383), including its two boundary quirks
template <typename S, typename R>
void div(image3d<S>& res, std::vector<image3d<R> > const& fun)
{
    if (fun.size() != 3) {
        tk::warning("div: a 3-D vector field is expected");
        return;
    }
    res.shape_like(fun[0]);
    tk::check(tk_div(res.ctx(), fun[0].handle(), fun[1].handle(), fun[2].handle(), res.handle()),
              "div");
}

}  // namespace im3d

#endif
