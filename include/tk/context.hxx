/*
 * tk/context.hxx -- device context and error handling of the C++ class layer.
 *
 * The reference is a host library with no device or stream concept; its classes are
 * used from one host thread.  The class layer keeps that usage: every object works on
 * the calling thread's current context (device 0 unless tk::context::use() picked
 * another), which owns one HIP stream, so all work of a thread is ordered.
 *
 * Errors: the reference prints "WARNING::..." and keeps its previous state for bad
 * arguments (e.g. rsfe_splitbregman_imp.hxx:1030-1047); the class layer does the same for
 * argument errors the reference checks, and throws tk::error for failures the reference
 * has no counterpart for (device errors, allocation failures).
 */
#ifndef TK_CONTEXT_HXX
#define TK_CONTEXT_HXX

#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "../tk.h"

namespace tk {

class error : public std::runtime_error {
public:
    error(int status, const std::string& what)
        : std::runtime_error(what + ": " + tk_last_error()), status_(status)
    {
    }
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char* what)
{
    if (status != TK_OK) throw error(status, what);
}

inline void warning(const std::string& msg) { std::clog << "WARNING::" << msg << std::endl; }

// element type of a volume
template <class T>
struct dtype_of;
template <>
struct dtype_of<double> {
    static constexpr tk_dtype value = TK_F64;
};
template <>
struct dtype_of<float> {
    static constexpr tk_dtype value = TK_F32;
};
template <>
struct dtype_of<unsigned char> {
    static constexpr tk_dtype value = TK_U8;
};

class context {
public:
    explicit context(int device = 0)
    {
        check(tk_ctx_create(device, &h_), "tk_ctx_create");
    }
    ~context()
    {
        if (h_) tk_ctx_destroy(h_);
    }
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    tk_ctx handle() const { return h_; }
    void sync() const { check(tk_ctx_sync(h_), "tk_ctx_sync"); }

    // the calling thread's current context (created on first use, device 0)
    static context& current()
    {
        std::shared_ptr<context>& c = slot();
        if (!c) c = std::make_shared<context>(0);
        return *c;
    }
    // make a context on `device` current for the calling thread; objects created before
    // keep the context they were created on
    static void use(int device) { slot() = std::make_shared<context>(device); }
    static std::shared_ptr<context> current_ptr()
    {
        current();
        return slot();
    }

private:
    static std::shared_ptr<context>& slot()
    {
        thread_local std::shared_ptr<context> c;
        return c;
    }
    tk_ctx h_ = nullptr;
};

}  // namespace tk

#endif
