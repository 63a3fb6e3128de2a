// A 3D real array resident in HBM with the reference's padded layout (C order, k fastest, every
// index starting at -1: lib/field/field.cc:90-91).  It stands where the drop-in classes hold a
// Blitz array: the data live on the device, the host sees them through explicit copies
// (upload / download), and a failed device call prints and exits as the reference's fatal
// conditions do.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "saras_mg.h"
#include "sr_array.h"

class DArray3 {
  public:
    DArray3() = default;
    DArray3(int nx, int ny, int nz) { resize(nx, ny, nz); }
    ~DArray3() { release(); }
    DArray3(const DArray3 &) = delete;
    DArray3 &operator=(const DArray3 &) = delete;
    DArray3(DArray3 &&o) noexcept { swap(o); }
    DArray3 &operator=(DArray3 &&o) noexcept {
        swap(o);
        return *this;
    }

    // core sizes nx, ny, nz; (nx + 2)(ny + 2)(nz + 2) entries, zero-filled
    void resize(int nx, int ny, int nz) {
        release();
        nx_ = nx; ny_ = ny; nz_ = nz;
        void *p = nullptr;
        check(sr_device_alloc((unsigned long)(bytes()), &p));
        d_ = (real *)p;
        *this = 0.0;
    }
    real *data() { return d_; }                 // device pointer
    const real *data() const { return d_; }
    size_t numElements() const { return (size_t)(nx_ + 2) * (ny_ + 2) * (nz_ + 2); }
    int lbound(int) const { return -1; }
    int ubound(int d) const { return d == 0 ? nx_ : d == 1 ? ny_ : nz_; }

    void upload(const Array3 &h) {
        if (h.numElements() != numElements()) fail("upload: shapes differ");
        check(sr_memcpy(d_, h.data(), (unsigned long)bytes(), 1));
    }
    void download(Array3 &h) const {
        h.resize(nx_, ny_, nz_);
        check(sr_memcpy(h.data(), d_, (unsigned long)bytes(), 2));
    }
    Array3 host() const {
        Array3 h;
        download(h);
        return h;
    }
    // every entry, pads included (Blitz's `F = a`)
    DArray3 &operator=(real a) {
        std::vector<real> v(numElements(), a);
        check(sr_memcpy(d_, v.data(), (unsigned long)bytes(), 1));
        return *this;
    }
    void copyFrom(const DArray3 &o) {
        if (o.numElements() != numElements()) fail("copy: shapes differ");
        check(sr_memcpy(d_, o.d_, (unsigned long)bytes(), 3));
    }

  private:
    size_t bytes() const { return numElements() * sizeof(real); }
    void release() {
        if (d_) sr_device_free(d_);
        d_ = nullptr;
    }
    void swap(DArray3 &o) {
        std::swap(d_, o.d_);
        std::swap(nx_, o.nx_);
        std::swap(ny_, o.ny_);
        std::swap(nz_, o.nz_);
    }
    static void fail(const char *what) {
        std::printf("DArray3 %s. ABORTING\n", what);
        std::exit(EXIT_FAILURE);
    }
    static void check(int status) {
        if (status != SR_OK) {
            std::printf("Device array error %d: %s. ABORTING\n", status, sr_last_error());
            std::exit(EXIT_FAILURE);
        }
    }
    real *d_ = nullptr;
    int nx_ = 0, ny_ = 0, nz_ = 0;
};
