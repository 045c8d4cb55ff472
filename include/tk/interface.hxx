/*
 * tk/interface.hxx -- im3d::interface<S> for MetaImage files (the reference's dump format).
 *
 * Mirrors the file part of im3d::interface<S> (header/interface.hxx:94-128, 279-280): the
 * constructor from a file name (interface_imp.hxx:33-88), the constructor from an image and
 * convertfromimage3d (97-122: the voxels as double with the image's spacing),
 * convert2image3d (125-150: cast back to S) and write(imagename, extension = ".mhd")
 * (672-711).  Only MetaImage (.mhd + .raw) is built: the reference's DICOM, JPEG and BMP
 * readers, the JPEG writer and every display member need VTK and are out of scope; a
 * constructor from another file type prints a WARNING and holds no image, write with
 * another extension prints the reference's ERROR line.  The image is held on the device
 * (libtk's tk_vol_read_mhd / tk_vol_write_mhd do the file work and the voxel reordering).
 */
#ifndef TK_INTERFACE_HXX
#define TK_INTERFACE_HXX

#include <iostream>
#include <string>

#include "image3d.hxx"

namespace im3d {

template <typename S>
class interface {
public:
    interface() : ctx_(tk::context::current_ptr()) {}

    explicit interface(std::string const& imagename) : ctx_(tk::context::current_ptr())
    {
        if (imagename.size() >= 4 && imagename.compare(imagename.size() - 4, 4, ".mhd") == 0)
            tk::check(tk_vol_read_mhd(ctx_->handle(), imagename.c_str(), TK_F64, &v_),
                      "interface(file)");
        else
            tk::warning("interface: only MetaImage (.mhd) files are read by this build: " +
                        imagename);
    }

    explicit interface(image3d<S> const& myim) : ctx_(tk::context::current_ptr())
    {
        convertfromimage3d(myim);
    }

    interface(interface const&) = delete;
    interface& operator=(interface const&) = delete;
    virtual ~interface() { release(); }

    void convertfromimage3d(image3d<S> const& myim)
    {
        release();
        if (!myim.handle()) return;
        tk::check(tk_vol_alloc(ctx_->handle(), (int)myim.getdimx(), (int)myim.getdimy(),
                               (int)myim.getdimz(), myim.gethx(), myim.gethy(), myim.gethz(),
                               TK_F64, &v_),
                  "convertfromimage3d");
        tk::check(tk_vol_copy(v_, myim.handle()), "convertfromimage3d");
    }

    void convert2image3d(image3d<S>& myim) const
    {
        if (!v_) {
            myim = image3d<S>();
            return;
        }
        int d[3];
        double h[3];
        tk_dtype dt;
        tk::check(tk_vol_info(v_, d, h, &dt), "convert2image3d");
        myim = image3d<S>((uint)d[0], (uint)d[1], (uint)d[2], h[0], h[1], h[2]);
        tk::check(tk_vol_copy(myim.handle(), v_), "convert2image3d");
    }

    void write(std::string const& imagename = "image3d",
               std::string const& extension = ".mhd") const
    {
        if (extension != ".mhd") {
            std::cout << "ERROR: " << imagename << " has an extension " << extension
                      << " not allowed in im3d::interface<S>::write" << std::endl;
            return;
        }
        if (!v_) {
            tk::warning("interface::write: no image");
            return;
        }
        tk::check(tk_vol_write_mhd(v_, imagename.c_str()), "interface::write");
    }

private:
    void release()
    {
        if (v_) tk_vol_free(v_);
        v_ = nullptr;
    }

    std::shared_ptr<tk::context> ctx_;
    tk_vol v_ = nullptr;
};

}  // namespace im3d

#endif
