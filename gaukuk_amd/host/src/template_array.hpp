// template_array.hpp -- owning, zero-initialised, SoA host array with the index convention of
// the reference's TArray (src/template_array.hpp:10-87): element (n, k, j, i) of an array of
// shape (n4, n3, n2, n1) lives at i + n1*(j + n2*(k + n3*n)), x fastest.
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace Gaukuk {

template <class T>
class TArray {
public:
    TArray() = default;
    explicit TArray(int n1) { NewArray(n1); }
    TArray(int n2, int n1) { NewArray(n2, n1); }
    TArray(int n3, int n2, int n1) { NewArray(n3, n2, n1); }
    TArray(int n4, int n3, int n2, int n1) { NewArray(n4, n3, n2, n1); }

    void NewArray(int n1) { Shape(1, 1, 1, n1); }
    void NewArray(int n2, int n1) { Shape(1, 1, n2, n1); }
    void NewArray(int n3, int n2, int n1) { Shape(1, n3, n2, n1); }
    void NewArray(int n4, int n3, int n2, int n1) { Shape(n4, n3, n2, n1); }

    size_t GetSize() const { return v_.size(); }
    size_t GetSizeInBytes() const { return v_.size() * sizeof(T); }
    int GetN1() const { return n1_; }
    int GetN2() const { return n2_; }
    int GetN3() const { return n3_; }
    int GetN4() const { return n4_; }

    T& operator()(int i) { return v_[i]; }
    T operator()(int i) const { return v_[i]; }
    T& operator()(int n, int i) { return v_[At(0, 0, n, i)]; }
    T operator()(int n, int i) const { return v_[At(0, 0, n, i)]; }
    T& operator()(int n, int j, int i) { return v_[At(0, n, j, i)]; }
    T operator()(int n, int j, int i) const { return v_[At(0, n, j, i)]; }
    T& operator()(int n, int k, int j, int i) { return v_[At(n, k, j, i)]; }
    T operator()(int n, int k, int j, int i) const { return v_[At(n, k, j, i)]; }

    T* data() noexcept { return v_.data(); }
    const T* data() const noexcept { return v_.data(); }
    void Swap(TArray& o) noexcept {
        std::swap(v_, o.v_);
        std::swap(n1_, o.n1_); std::swap(n2_, o.n2_); std::swap(n3_, o.n3_); std::swap(n4_, o.n4_);
    }

private:
    void Shape(int n4, int n3, int n2, int n1) {
        n1_ = n1; n2_ = n2; n3_ = n3; n4_ = n4;
        v_.assign((size_t)n1 * n2 * n3 * n4, T());
    }
    size_t At(int n, int k, int j, int i) const {
        return (size_t)i + (size_t)n1_ * ((size_t)j + (size_t)n2_ * ((size_t)k + (size_t)n3_ * n));
    }
    std::vector<T> v_;
    int n1_ = 0, n2_ = 0, n3_ = 0, n4_ = 0;
};

}  // namespace Gaukuk
