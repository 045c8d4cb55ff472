/*
 * tk/poisson.hxx -- lapl::unsteady_poisson_functor<T>, lapl::DirGaussSeidel and
 * lapl::NeuGaussSeidel over the device Gauss-Seidel sweeps of libtk.
 *
 * Mirrors lapl::unsteady_poisson_functor<T> (header/poisson.hxx:54-111): a functor holding
 * void (*f)(image3d<T>& res, image3d<T> const& b, T const& dt, T const& bc,
 * image3d<T> const& input) (60-61), its constructor (69-71), operator() with defaults
 * dt = 1, bc = 0, input = empty image (90-96) and changef (104-110).  One call is one sweep,
 * in place on res; input is unused, as in both reference built-ins.
 *
 * DirGaussSeidel (poisson.hxx:153; poisson_imp.hxx:340-429) sets the six boundary faces of
 * res to bc and runs one lexicographic sweep: on the device this is a pipelined wavefront
 * whose result is bitwise the reference's sequential order.  DirRedBlackGaussSeidel (no
 * reference counterpart) is the red-black ordering of the same update: the same fixed
 * point, other iterates.  NeuGaussSeidel (poisson.hxx:193; poisson_imp.hxx:115-334) updates
 * every voxel with the derivative bc prescribed on the boundary (3-D branch; bitwise the
 * reference in FP64).  Like the reference, res must already have b's dims; a mismatch
 * prints a WARNING and leaves res unchanged (the reference does not check at all).
 */
#ifndef TK_POISSON_HXX
#define TK_POISSON_HXX

#include "image3d.hxx"

namespace lapl {

template <typename T>
class unsteady_poisson_functor {
public:
    typedef void (*function)(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt,
                             T const& bc, im3d::image3d<T> const& input);

    explicit unsteady_poisson_functor(function f) : f_(f) {}

    void operator()(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt = 1.,
                    T const& bc = 0., im3d::image3d<T> const& input = im3d::image3d<T>())
    {
        f_(res, b, dt, bc, input);
    }

    void changef(function f) { f_ = f; }
    function getf() const { return f_; }

private:
    function f_;
};

namespace detail {
template <typename T>
void dir_sweep(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt, T const& bc,
               int order, const char* name, bool neumann = false)
{
    if (res.getdimx() != b.getdimx() || res.getdimy() != b.getdimy() ||
        res.getdimz() != b.getdimz()) {
        tk::warning(std::string(name) + ": res and b must have the same dimensions");
        return;
    }
    if (neumann)
        tk::check(tk_poisson_neu_sweeps(res.ctx(), res.handle(), b.handle(), (double)dt,
                                        (double)bc, order, 1),
                  name);
    else
        tk::check(tk_poisson_dir_sweeps(res.ctx(), res.handle(), b.handle(), (double)dt,
                                        (double)bc, order, 1),
                  name);
}
}  // namespace detail

// one Dirichlet Gauss-Seidel sweep, lexicographic (poisson_imp.hxx:340-429)
template <typename T>
void DirGaussSeidel(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt, T const& bc,
                    im3d::image3d<T> const& input)
{
    (void)input;
    detail::dir_sweep(res, b, dt, bc, TK_GS_LEX, "DirGaussSeidel");
}

// one Neumann Gauss-Seidel sweep, lexicographic (poisson_imp.hxx:115-334)
template <typename T>
void NeuGaussSeidel(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt, T const& bc,
                    im3d::image3d<T> const& input)
{
    (void)input;
    detail::dir_sweep(res, b, dt, bc, TK_GS_LEX, "NeuGaussSeidel", true);
}

// one Dirichlet Gauss-Seidel sweep, red-black order (build addition)
template <typename T>
void DirRedBlackGaussSeidel(im3d::image3d<T>& res, im3d::image3d<T> const& b, T const& dt,
                            T const& bc, im3d::image3d<T> const& input)
{
    (void)input;
    detail::dir_sweep(res, b, dt, bc, TK_GS_REDBLACK, "DirRedBlackGaussSeidel");
}

}  // namespace lapl

#endif
