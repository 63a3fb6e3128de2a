// Minimal stand-in for the Blitz++ arrays the SARAS interfaces pass around: a 3D real array in
// C order (k fastest) whose every index starts at -1, the layout of every padded field of the
// reference (lib/field/field.cc:90-91, lib/poisson/poisson.cc:300-301), and a 1D array likewise.
#pragma once

#include <cstddef>
#include <vector>

typedef double real;

class Array1 {
  public:
    Array1() = default;
    explicit Array1(int n) : n_(n), v_((size_t)n + 2, 0.0) {}   // indices -1..n
    real &operator()(int i) { return v_[(size_t)(i + 1)]; }
    real operator()(int i) const { return v_[(size_t)(i + 1)]; }
    real *data() { return v_.data(); }
    const real *data() const { return v_.data(); }
    int size() const { return n_; }

  private:
    int n_ = 0;
    std::vector<real> v_;
};

class Array3 {
  public:
    Array3() = default;
    Array3(int nx, int ny, int nz) { resize(nx, ny, nz); }
    void resize(int nx, int ny, int nz) {
        nx_ = nx; ny_ = ny; nz_ = nz;
        v_.assign((size_t)(nx + 2) * (ny + 2) * (nz + 2), 0.0);
    }
    real &operator()(int i, int j, int k) { return v_[idx(i, j, k)]; }
    real operator()(int i, int j, int k) const { return v_[idx(i, j, k)]; }
    Array3 &operator=(real a) {
        for (auto &x : v_) x = a;
        return *this;
    }
    real *data() { return v_.data(); }
    const real *data() const { return v_.data(); }
    size_t numElements() const { return v_.size(); }
    int lbound(int) const { return -1; }
    int ubound(int d) const { return (d == 0 ? nx_ : d == 1 ? ny_ : nz_); }

  private:
    size_t idx(int i, int j, int k) const {
        return ((size_t)(i + 1) * (ny_ + 2) + (j + 1)) * (nz_ + 2) + (k + 1);
    }
    int nx_ = 0, ny_ = 0, nz_ = 0;
    std::vector<real> v_;
};
