// blocksp.hpp -- C++20 drop-in for BlockSp's block-sparse solve path on MI355X.
//
// Mirrors the reference's template surface (namespace BlockSp):
//   * containers  BlockSp<T,B> = Sp::CSR<TinyMatrix<T,B,B>>, solArray<T,B,D>,
//                 dot / twoNorm / infNorm / axpy            (blocksp_containers.hpp)
//   * products    bsmv (4 forms)                             (blocksp_multiply.hpp:365-390)
//   * preconditioners NoPre, Diag<Inv>, SOR<Inv>, ILU0, IChol0 (blocksp_preconditioners.hpp)
//   * solvers     pCG, pGMRES (+ deflated restart), SOR_pGMRES, SOR_pCG, ILU0_pGMRES,
//                 IChol0_pCG, ez::*                          (blocksp_solvers.hpp)
// The containers live on the host exactly as in the reference (insertion-ordered
// rows); every product, preconditioner application and Krylov iteration runs on
// the GPU through the C ABI in blocksp_c.h.  A matrix is flattened to BSR and
// uploaded once and cached until it is modified (non-const access bumps a
// version stamp).  Device arithmetic is FP64: float matrices/vectors are widened
// on upload and narrowed on download.  Complex T (std::complex<float/double>) uploads through
// bsp_bsr_create_complex; its vectors travel as interleaved (re, im) doubles.
//
// Link with -lblocksp_amd (blocksp_amd/libblocksp_amd.so).
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../blocksp_c.h"

namespace BlockSp {

// ===========================================================================
// errors and type traits (blocksp_containers.hpp:59-83)
// ===========================================================================

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <typename T> struct is_complex : std::false_type {};
template <std::floating_point T> struct is_complex<std::complex<T>> : std::true_type {};
template <typename T> inline constexpr bool is_complex_v = is_complex<T>::value;
template <typename T> struct ScalarType { using type = T; };
template <typename U> struct ScalarType<std::complex<U>> { using type = U; };
template <typename T> using ScalarType_t = typename ScalarType<T>::type;

// ===========================================================================
// small fixed-size vectors and matrices (the blitz::Tiny* roles)
// ===========================================================================

template <typename T, int N>
struct TinyVector {
    std::array<T, N> v{};
    TinyVector() = default;
    explicit TinyVector(T s) { v.fill(s); }
    template <typename... Ts>
        requires(sizeof...(Ts) == N && N > 1)
    TinyVector(Ts... xs) : v{static_cast<T>(xs)...} {}
    T& operator()(int i) { return v[i]; }
    const T& operator()(int i) const { return v[i]; }
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    T* data() { return v.data(); }
    const T* data() const { return v.data(); }
    static constexpr int length() { return N; }
    TinyVector& operator=(T s) { v.fill(s); return *this; }
    TinyVector& operator+=(const TinyVector& o) { for (int i = 0; i < N; ++i) v[i] += o.v[i]; return *this; }
    TinyVector& operator-=(const TinyVector& o) { for (int i = 0; i < N; ++i) v[i] -= o.v[i]; return *this; }
    TinyVector& operator+=(T s) { for (auto& e : v) e += s; return *this; }
    TinyVector& operator-=(T s) { for (auto& e : v) e -= s; return *this; }
    TinyVector& operator*=(T s) { for (auto& e : v) e *= s; return *this; }
    TinyVector& operator/=(T s) { for (auto& e : v) e /= s; return *this; }
    friend TinyVector operator+(TinyVector a, const TinyVector& b) { return a += b; }
    friend TinyVector operator-(TinyVector a, const TinyVector& b) { return a -= b; }
    friend TinyVector operator-(TinyVector a) { for (auto& e : a.v) e = -e; return a; }
    friend TinyVector operator*(TinyVector a, T s) { return a *= s; }
    friend TinyVector operator*(T s, TinyVector a) { return a *= s; }
    friend TinyVector operator/(TinyVector a, T s) { return a /= s; }
    friend bool operator==(const TinyVector&, const TinyVector&) = default;
};

template <typename T, int M, int N>
struct TinyMatrix {
    std::array<T, M * N> a{};
    TinyMatrix() = default;
    explicit TinyMatrix(T s) { a.fill(s); }
    T& operator()(int i, int j) { return a[i * N + j]; }
    const T& operator()(int i, int j) const { return a[i * N + j]; }
    T* data() { return a.data(); }
    const T* data() const { return a.data(); }
    TinyMatrix& operator=(T s) { a.fill(s); return *this; }
    TinyMatrix& operator+=(const TinyMatrix& o) { for (int i = 0; i < M * N; ++i) a[i] += o.a[i]; return *this; }
    TinyMatrix& operator-=(const TinyMatrix& o) { for (int i = 0; i < M * N; ++i) a[i] -= o.a[i]; return *this; }
    TinyMatrix& operator*=(T s) { for (auto& e : a) e *= s; return *this; }
    TinyMatrix& operator/=(T s) { for (auto& e : a) e /= s; return *this; }
    friend TinyMatrix operator+(TinyMatrix x, const TinyMatrix& y) { return x += y; }
    friend TinyMatrix operator-(TinyMatrix x, const TinyMatrix& y) { return x -= y; }
    friend TinyMatrix operator*(TinyMatrix x, T s) { return x *= s; }
    friend TinyMatrix operator*(T s, TinyMatrix x) { return x *= s; }
    friend bool operator==(const TinyMatrix&, const TinyMatrix&) = default;
};

namespace detail {
inline std::uint64_t next_id() {
    static std::atomic<std::uint64_t> id{1};
    return id.fetch_add(1);
}
// called when a matrix is destroyed so that the device cache can drop it
inline void (*&release_hook())(std::uint64_t) {
    static void (*hook)(std::uint64_t) = nullptr;
    return hook;
}
}  // namespace detail

// ===========================================================================
// Sp::CSR (blocksp_containers.hpp:92-496): rows keep insertion order
// ===========================================================================

namespace Sp {

template <typename T>
class CSR {
    int m_m = 0, m_n = 0, nnz = 0;
    T m_zero{};
    std::vector<std::vector<T>> m_data;
    std::vector<std::vector<int>> m_colIndex;
    std::uint64_t m_id = detail::next_id();
    mutable std::uint64_t m_version = 0;

    void touch() const { ++m_version; }

public:
    CSR() { clear_zero(); }
    CSR(int m, int n) : m_m(m), m_n(n), m_data(m), m_colIndex(m) { clear_zero(); }
    explicit CSR(int n) : CSR(n, n) {}
    CSR(const CSR& o)
        : m_m(o.m_m), m_n(o.m_n), nnz(o.nnz), m_zero(o.m_zero), m_data(o.m_data), m_colIndex(o.m_colIndex) {}
    CSR& operator=(const CSR& o) {
        m_m = o.m_m; m_n = o.m_n; nnz = o.nnz; m_zero = o.m_zero;
        m_data = o.m_data; m_colIndex = o.m_colIndex;
        touch();
        return *this;
    }
    ~CSR() {
        if (detail::release_hook()) detail::release_hook()(m_id);
    }
    void clear_zero() {
        if constexpr (requires(T t) { t = 0.0; }) m_zero = T(0.0);
    }

    // Non-const indexing inserts a zero block when (i,j) is absent (containers.hpp:147-164).
    T& operator()(int i, int j) {
        assert(0 <= i && i < m_m && 0 <= j && j < m_n);
        touch();
        auto& ci = m_colIndex[i];
        for (size_t in = 0; in < ci.size(); ++in)
            if (ci[in] == j) return m_data[i][in];
        ci.push_back(j);
        m_data[i].push_back(m_zero);
        ++nnz;
        return m_data[i].back();
    }
    const T& operator()(int i, int j) const {
        const auto& ci = m_colIndex[i];
        for (size_t in = 0; in < ci.size(); ++in)
            if (ci[in] == j) return m_data[i][in];
        return m_zero;
    }
    T& at(int i, int j) {
        touch();
        auto& ci = m_colIndex[i];
        for (size_t in = 0; in < ci.size(); ++in)
            if (ci[in] == j) return m_data[i][in];
        return m_zero;
    }
    const T& at(int i, int j) const { return (*this)(i, j); }
    bool scan(int i, int j) const {
        for (int c : m_colIndex[i])
            if (c == j) return true;
        return false;
    }
    const std::vector<std::vector<T>>& get_data() const { return m_data; }
    std::vector<std::vector<T>>& get_data() { touch(); return m_data; }
    const std::vector<T>& get_data(int im) const { return m_data[im]; }
    std::vector<T>& get_data(int im) { touch(); return m_data[im]; }
    const std::vector<std::vector<int>>& get_colIndex() const { return m_colIndex; }
    const std::vector<int>& get_colIndex(int im) const { return m_colIndex[im]; }
    TinyVector<int, 2> size() const { return TinyVector<int, 2>(m_m, m_n); }
    int size(int dim) const { return dim == 0 ? m_m : m_n; }
    int m() const { return m_m; }
    int n() const { return m_n; }
    int nz() const { return nnz; }
    size_t colSize(int im) const { return m_colIndex[im].size(); }
    int colInd(int im, int in) const { return m_colIndex[im][in]; }
    const std::vector<int>& colInd(int im) const { return m_colIndex[im]; }
    void addCol(int nCol = 1) { m_n += nCol; touch(); }
    // lowTri (containers.hpp:277-285)
    CSR lowTri(int i = 0) const {
        CSR low(m_m, m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < m_colIndex[im].size(); ++in)
                if (m_colIndex[im][in] - i <= im) low(im, m_colIndex[im][in]) = m_data[im][in];
        return low;
    }
    // upTri (containers.hpp:287-296)
    CSR upTri(int i = 0) const {
        CSR up(m_m, m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < m_colIndex[im].size(); ++in)
                if (m_colIndex[im][in] - i >= im) up(im, m_colIndex[im][in]) = m_data[im][in];
        return up;
    }
    // transpose (containers.hpp:264-274): the structure is transposed and each element copied
    // as it is (conjugated for complex scalars); bs_tr also transposes the blocks.
    CSR transpose() const {
        CSR tr(m_n, m_m);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < m_colIndex[im].size(); ++in) {
                if constexpr (is_complex_v<T>) tr(m_colIndex[im][in], im) = std::conj(m_data[im][in]);
                else tr(m_colIndex[im][in], im) = m_data[im][in];
            }
        return tr;
    }
    // arithmetic (containers.hpp:302-353): blocks of the other matrix are added in its row
    // order, inserting the ones this matrix lacks at the end of the row
    CSR& operator+=(const CSR& o) {
        assert(o.m_m == m_m && o.m_n == m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < o.m_colIndex[im].size(); ++in) (*this)(im, o.m_colIndex[im][in]) += o.m_data[im][in];
        return *this;
    }
    CSR& operator-=(const CSR& o) {
        assert(o.m_m == m_m && o.m_n == m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < o.m_colIndex[im].size(); ++in) (*this)(im, o.m_colIndex[im][in]) -= o.m_data[im][in];
        return *this;
    }
    CSR operator+(const CSR& o) const { CSR r = *this; return r += o; }
    CSR operator-(const CSR& o) const { CSR r = *this; return r -= o; }
    // scaling by a scalar (or, for scalar elements, by an element)
    template <typename S>
        requires requires(T t, S s) { t *= s; }
    CSR& operator*=(const S& s) {
        touch();
        for (auto& row : m_data)
            for (auto& e : row) e *= s;
        return *this;
    }
    template <typename S>
        requires requires(T t, S s) { t /= s; }
    CSR& operator/=(const S& s) {
        touch();
        for (auto& row : m_data)
            for (auto& e : row) e /= s;
        return *this;
    }
    // addLarger / subtractLarger (containers.hpp:359-378): rows of a taller matrix, my rows only
    CSR& addLarger(const CSR& o) {
        assert(o.m_m >= m_m && o.m_n == m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < o.m_colIndex[im].size(); ++in) (*this)(im, o.m_colIndex[im][in]) += o.m_data[im][in];
        return *this;
    }
    CSR& subtractLarger(const CSR& o) {
        assert(o.m_m >= m_m && o.m_n == m_n);
        for (int im = 0; im < m_m; ++im)
            for (size_t in = 0; in < o.m_colIndex[im].size(); ++in) (*this)(im, o.m_colIndex[im][in]) -= o.m_data[im][in];
        return *this;
    }
    // rowIndicesByCol (containers.hpp:410-417)
    std::vector<std::unordered_set<int>> rowIndicesByCol() const {
        std::vector<std::unordered_set<int>> ind(m_n);
        for (int im = 0; im < m_m; ++im)
            for (int c : m_colIndex[im]) ind[c].emplace(im);
        return ind;
    }
    std::vector<TinyVector<int, 2>> getIndices() const {
        std::vector<TinyVector<int, 2>> out;
        for (int im = 0; im < m_m; ++im)
            for (int c : m_colIndex[im]) out.emplace_back(im, c);
        return out;
    }
    void shrink_to_fit() {
        for (int im = 0; im < m_m; ++im) {
            m_colIndex[im].shrink_to_fit();
            m_data[im].shrink_to_fit();
        }
    }
    void reserve_col(int im, int s) { m_colIndex[im].reserve(s); }
    void reserve_data(int im, int s) { m_data[im].reserve(s); }
    void reserve(int im, int s) { reserve_col(im, s); reserve_data(im, s); }
    std::uint64_t id() const { return m_id; }
    std::uint64_t version() const { return m_version; }
};

/// Scalar products of a sparse matrix (containers.hpp:498-504).
template <typename T, typename S>
    requires requires(T t, S s) { t *= s; }
CSR<T> operator*(const S& s, CSR<T> A) { return A *= s; }
template <typename T, typename S>
    requires requires(T t, S s) { t *= s; }
CSR<T> operator*(CSR<T> A, const S& s) { return A *= s; }
template <typename T, typename S>
    requires requires(T t, S s) { t /= s; }
CSR<T> operator/(CSR<T> A, const S& s) { return A /= s; }

}  // namespace Sp

/// BlockSp<T,B>: a Sp::CSR matrix of dense B x B blocks (blocksp_containers.hpp:52-53).
template <typename T, int B>
using BlockSp = Sp::CSR<TinyMatrix<T, B, B>>;

// ===========================================================================
// solArray<T,B,D> (blocksp_containers.hpp:572-736): flat (ie, dim, ib) layout
// ===========================================================================

template <typename T, int B, int D = 1>
class solArray {
    using Elem = TinyVector<TinyVector<T, B>, D>;
    static_assert(sizeof(Elem) == sizeof(T) * B * D, "solArray elements must be densely packed");
    std::vector<Elem> m_arr;
    int m_len = 0;

public:
    solArray() = default;
    explicit solArray(int len) : m_arr(len), m_len(len) { set_to_zero(); }
    solArray(int len, T val) : m_arr(len), m_len(len) { fill(val); }

    int len() const { return m_len; }
    void resize(int l, T val = T(0)) { m_len = l; m_arr.assign(l, Elem{}); fill(val); }
    void resize_and_preserve(int l) { m_len = l; m_arr.resize(l); }
    void reset() { m_len = 0; m_arr.clear(); }
    void set_to_zero() { fill(T(0)); }
    void fill(T val) { std::fill(data(), data() + size_flat(), val); }
    T* data() { return reinterpret_cast<T*>(m_arr.data()); }
    const T* data() const { return reinterpret_cast<const T*>(m_arr.data()); }
    size_t size_flat() const { return (size_t)m_len * B * D; }

    template <int DD = D> requires(DD == 1)
    TinyVector<T, B>& operator()(int im, int = 0) { return m_arr[im](0); }
    template <int DD = D> requires(DD == 1)
    const TinyVector<T, B>& operator()(int im, int = 0) const { return m_arr[im](0); }
    template <int DD = D> requires(DD > 1)
    Elem& operator()(int im) { return m_arr[im]; }
    template <int DD = D> requires(DD > 1)
    const Elem& operator()(int im) const { return m_arr[im]; }
    template <int DD = D> requires(DD > 1)
    TinyVector<T, B>& operator()(int im, int dim) { return m_arr[im](dim); }
    template <int DD = D> requires(DD > 1)
    const TinyVector<T, B>& operator()(int im, int dim) const { return m_arr[im](dim); }

    solArray& operator+=(T val) { for (size_t i = 0; i < size_flat(); ++i) data()[i] += val; return *this; }
    solArray& operator-=(T val) { for (size_t i = 0; i < size_flat(); ++i) data()[i] -= val; return *this; }
    solArray& operator*=(T val) { for (size_t i = 0; i < size_flat(); ++i) data()[i] *= val; return *this; }
    solArray& operator/=(T val) { for (size_t i = 0; i < size_flat(); ++i) data()[i] /= val; return *this; }
    solArray& operator+=(const solArray& o) { for (size_t i = 0; i < size_flat(); ++i) data()[i] += o.data()[i]; return *this; }
    solArray& operator-=(const solArray& o) { for (size_t i = 0; i < size_flat(); ++i) data()[i] -= o.data()[i]; return *this; }
    solArray& operator*=(const solArray& o) { for (size_t i = 0; i < size_flat(); ++i) data()[i] *= o.data()[i]; return *this; }
    solArray& operator/=(const solArray& o) { for (size_t i = 0; i < size_flat(); ++i) data()[i] /= o.data()[i]; return *this; }

    auto begin() { return m_arr.begin(); }
    auto end() { return m_arr.end(); }
    auto begin() const { return m_arr.begin(); }
    auto end() const { return m_arr.end(); }
};

template <typename T, int B, int D>
solArray<T, B, D> operator+(solArray<T, B, D> a, const solArray<T, B, D>& b) { return a += b; }
template <typename T, int B, int D>
solArray<T, B, D> operator-(solArray<T, B, D> a, const solArray<T, B, D>& b) { return a -= b; }
template <typename T, int B, int D>
solArray<T, B, D> operator*(solArray<T, B, D> a, T s) { return a *= s; }
template <typename T, int B, int D>
solArray<T, B, D> operator*(T s, solArray<T, B, D> a) { return a *= s; }
template <typename T, int B, int D>
solArray<T, B, D> operator/(solArray<T, B, D> a, T s) { return a /= s; }

/// y += a x (containers.hpp:857-864)
template <typename T, int B, int D>
inline void axpy(solArray<T, B, D>& y, T a, const solArray<T, B, D>& x) {
    assert(x.len() == y.len());
    for (size_t i = 0; i < x.size_flat(); ++i) y.data()[i] += a * x.data()[i];
}

/// dot (containers.hpp:974-996): running sum of per-block dot products.
template <typename T, int B, int D = 1>
T dot(const solArray<T, B, D>& x1, const solArray<T, B, D>& x2, int dim = -1) {
    assert(x1.len() == x2.len());
    T d = T(0);
    for (int ie = 0; ie < x1.len(); ++ie)
        for (int dd = 0; dd < D; ++dd) {
            if (dim >= 0 && dd != dim) continue;
            const T* a = x1.data() + ((size_t)ie * D + dd) * B;
            const T* b = x2.data() + ((size_t)ie * D + dd) * B;
            T bd = a[0] * b[0];
            for (int i = 1; i < B; ++i) bd += a[i] * b[i];
            d += bd;
        }
    return d;
}

template <typename T, int B, int D = 1>
inline ScalarType_t<T> twoNorm(const solArray<T, B, D>& x, const solArray<T, B, D>& y, int dim = -1) {
    return std::sqrt(std::abs(dot<T, B, D>(x, y, dim)));
}
template <typename T, int B, int D = 1>
inline ScalarType_t<T> twoNorm(const solArray<T, B, D>& x, int dim = -1) {
    return twoNorm<T, B, D>(x, x, dim);
}
template <typename T, int B, int D = 1>
inline ScalarType_t<T> infNorm(const solArray<T, B, D>& sa, int dim = -1) {
    ScalarType_t<T> nrm = 0;
    for (int ie = 0; ie < sa.len(); ++ie)
        for (int dd = 0; dd < D; ++dd) {
            if (dim >= 0 && dd != dim) continue;
            const T* a = sa.data() + ((size_t)ie * D + dd) * B;
            for (int i = 0; i < B; ++i) nrm = std::max<ScalarType_t<T>>(nrm, std::abs(a[i]));
        }
    return nrm;
}

namespace util {
/// tvSubset / tvConcat (blocksp_utility.hpp:242-265): the D>1 <-> B*D reshape.
template <typename T, int B, int D>
inline TinyVector<T, B> tvSubset(const TinyVector<T, B * D>& tv, int dim) {
    TinyVector<T, B> sub;
    for (int ib = 0; ib < B; ++ib) sub(ib) = tv(B * dim + ib);
    return sub;
}
template <typename T, int B, int D>
inline TinyVector<T, B * D> tvConcat(const TinyVector<TinyVector<T, B>, D>& tv) {
    TinyVector<T, B * D> cat;
    for (int dim = 0; dim < D; ++dim)
        for (int ib = 0; ib < B; ++ib) cat(dim * B + ib) = tv(dim)(ib);
    return cat;
}
/// trace (blocksp_utility.hpp:350-380).
template <typename T, int B>
inline T trace(const TinyMatrix<T, B, B>& A) {
    T t = T(0);
    for (int i = 0; i < B; ++i) t += A(i, i);
    return t;
}
template <typename T, int B>
inline T trace(const BlockSp<T, B>& A) {
    T t = T(0);
    const int end = std::min(A.m(), A.n());
    for (int ie = 0; ie < end; ++ie)
        for (int ib = 0; ib < B; ++ib) t += A(ie, ie)(ib, ib);
    return t;
}
/// frobeniusNorm (blocksp_utility.hpp:328-336, 383-388): sqrt(trace(A^* A)), the square root
/// of the sum of squared magnitudes of the stored entries.
template <typename T, int M, int N>
inline ScalarType_t<T> frobeniusNorm(const TinyMatrix<T, M, N>& A) {
    ScalarType_t<T> s = 0;
    for (const T& e : A.a) s += std::norm(e);
    return std::sqrt(s);
}
template <typename T, int B>
inline ScalarType_t<T> frobeniusNorm(const BlockSp<T, B>& A) {
    ScalarType_t<T> s = 0;
    for (int im = 0; im < A.m(); ++im)
        for (const auto& blk : A.get_data(im))
            for (const T& e : blk.a) s += std::norm(e);
    return std::sqrt(s);
}
/// setFromGlobal (blocksp_utility.hpp:614-639).
template <typename T, int B>
void setFromGlobal(BlockSp<T, B>& A, int i, int j, T val) {
    A(i / B, j / B)(i % B, j % B) = val;
}
template <typename T, int B, int D = 1>
void setFromGlobal(solArray<T, B, D>& x, int i, T val) {
    x.data()[i] = val;
}
}  // namespace util

// ===========================================================================
// device layer over the C ABI
// ===========================================================================

namespace gpu {

inline void check(bsp_ctx* ctx, int rc, const char* what) {
    if (rc < 0) throw Error(rc, std::string(what) + ": " + bsp_last_error(ctx));
}

/// Process-wide device context (device from BLOCKSP_DEVICE, default 0).
class Context {
    bsp_ctx* h_ = nullptr;

public:
    Context() {
        const char* d = std::getenv("BLOCKSP_DEVICE");
        int rc = bsp_ctx_create(d ? std::atoi(d) : 0, nullptr, &h_);
        if (rc != BSP_OK) throw Error(rc, "bsp_ctx_create failed (is a GPU visible?)");
    }
    ~Context() { bsp_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    bsp_ctx* get() const { return h_; }
    static Context& instance() {
        static Context c;
        return c;
    }
};

inline bsp_ctx* ctx() { return Context::instance().get(); }

/// Doubles that n values of T occupy on the device (2 per complex value).
template <typename T>
constexpr size_t doubles(size_t n) {
    return is_complex_v<T> ? 2 * n : n;
}

/// Device vector of doubles; complex values are stored as (re, im) pairs.
class DVec {
    double* p_ = nullptr;
    size_t n_ = 0;

public:
    DVec() = default;
    explicit DVec(size_t n) : n_(n) {
        void* p = nullptr;
        check(ctx(), bsp_malloc(ctx(), sizeof(double) * std::max<size_t>(n, 1), &p), "bsp_malloc");
        p_ = static_cast<double*>(p);
    }
    ~DVec() { if (p_) bsp_free(ctx(), p_); }
    DVec(const DVec&) = delete;
    DVec& operator=(const DVec&) = delete;
    DVec(DVec&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    double* get() const { return p_; }
    size_t size() const { return n_; }
    // std::complex<S> is laid out as S[2], so complex arrays move as their scalar parts
    template <typename T>
    void upload(const T* src_t) {
        using S = ScalarType_t<T>;
        const S* src = reinterpret_cast<const S*>(src_t);
        if constexpr (std::is_same_v<S, double>) {
            check(ctx(), bsp_memcpy_h2d(ctx(), p_, src, sizeof(double) * n_), "h2d");
        } else {
            std::vector<double> tmp(src, src + n_);
            check(ctx(), bsp_memcpy_h2d(ctx(), p_, tmp.data(), sizeof(double) * n_), "h2d");
        }
    }
    template <typename T>
    void download(T* dst_t) const {
        using S = ScalarType_t<T>;
        S* dst = reinterpret_cast<S*>(dst_t);
        if constexpr (std::is_same_v<S, double>) {
            check(ctx(), bsp_memcpy_d2h(ctx(), dst, p_, sizeof(double) * n_), "d2h");
        } else {
            std::vector<double> tmp(n_);
            check(ctx(), bsp_memcpy_d2h(ctx(), tmp.data(), p_, sizeof(double) * n_), "d2h");
            for (size_t i = 0; i < n_; ++i) dst[i] = static_cast<S>(tmp[i]);
        }
    }
};

/// BSR copy of a BlockSp matrix on the device.
struct DMatrix {
    bsp_bsr* h = nullptr;
    int n_rows = 0, n_cols = 0, bs = 0;
    ~DMatrix() { if (h) bsp_bsr_destroy(h); }
};

class MatrixCache {
    std::mutex mu_;
    std::map<std::uint64_t, std::pair<std::uint64_t, std::shared_ptr<DMatrix>>> map_;

public:
    static MatrixCache& instance() {
        static MatrixCache c;
        return c;
    }
    void drop(std::uint64_t id) {
        std::lock_guard<std::mutex> g(mu_);
        map_.erase(id);
    }
    template <typename T, int M>
    std::shared_ptr<DMatrix> get(const BlockSp<T, M>& A) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = map_.find(A.id());
        if (it != map_.end() && it->second.first == A.version()) return it->second.second;
        // flatten to BSR in insertion order
        std::vector<int32_t> rp(A.m() + 1, 0), ci;
        std::vector<double> vals;   // complex entries as (re, im) pairs
        ci.reserve(A.nz());
        vals.reserve(doubles<T>((size_t)A.nz() * M * M));
        for (int i = 0; i < A.m(); ++i) {
            const auto& row = A.get_data(i);
            const auto& cols = A.get_colIndex(i);
            for (size_t k = 0; k < cols.size(); ++k) {
                ci.push_back(cols[k]);
                for (int e = 0; e < M * M; ++e) {
                    const T v = row[k].data()[e];
                    if constexpr (is_complex_v<T>) {
                        vals.push_back(static_cast<double>(v.real()));
                        vals.push_back(static_cast<double>(v.imag()));
                    } else {
                        vals.push_back(static_cast<double>(v));
                    }
                }
            }
            rp[i + 1] = (int32_t)ci.size();
        }
        auto d = std::make_shared<DMatrix>();
        if constexpr (is_complex_v<T>)
            check(ctx(), bsp_bsr_create_complex(ctx(), A.m(), A.n(), M, rp.data(), ci.data(), vals.data(), &d->h),
                  "bsr_create_complex");
        else
            check(ctx(), bsp_bsr_create(ctx(), A.m(), A.n(), M, rp.data(), ci.data(), vals.data(), &d->h),
                  "bsr_create");
        d->n_rows = A.m();
        d->n_cols = A.n();
        d->bs = M;
        map_[A.id()] = {A.version(), d};
        if (!detail::release_hook()) detail::release_hook() = [](std::uint64_t id) { MatrixCache::instance().drop(id); };
        return d;
    }
};

template <typename T, int M>
std::shared_ptr<DMatrix> device_matrix(const BlockSp<T, M>& A) {
    return MatrixCache::instance().get(A);
}

/// Owning handle of a device preconditioner.
struct DPrecond {
    bsp_precond* h = nullptr;
    std::shared_ptr<DMatrix> A;   // keeps the matrix alive (SOR/Diag reference it)
    ~DPrecond() { if (h) bsp_precond_destroy(h); }
};

}  // namespace gpu

// ===========================================================================
// bsmv (blocksp_multiply.hpp:365-390)
// ===========================================================================

namespace mult_detail {

template <typename T, int M, int B, int D>
void bsmv_device(solArray<T, B, D>& b, const BlockSp<T, M>& A, const solArray<T, B, D>& x, bool accumulate,
                 bool transpose, int dim_mode, int dim) {
    auto dA = gpu::device_matrix(A);
    gpu::DVec dx(gpu::doubles<T>(x.size_flat())), dy(gpu::doubles<T>(b.size_flat()));
    dx.upload(x.data());
    if (accumulate) dy.upload(b.data());
    int rc;
    if (dim_mode && transpose) {
        rc = bsp_bsmv_dim_transpose(gpu::ctx(), dA->h, D, dim, dx.get(), dy.get(), accumulate ? 1 : 0);
    } else if (dim_mode) {
        rc = bsp_bsmv_dim(gpu::ctx(), dA->h, D, dim, dx.get(), dy.get(), accumulate ? 1 : 0);
    } else if (transpose) {
        rc = bsp_bsmv_transpose(gpu::ctx(), dA->h, dx.get(), dy.get(), accumulate ? 1 : 0);
    } else {
        rc = bsp_bsmv(gpu::ctx(), dA->h, dx.get(), dy.get(), accumulate ? 1 : 0);
    }
    gpu::check(gpu::ctx(), rc, "bsmv");
    dy.download(b.data());
}

}  // namespace mult_detail

/// BlockSp<T,B> x solArray<T,B,D>, per dimension (dim < 0: all dimensions).
template <typename T, int B, int D = 1>
solArray<T, B, D> bsmv(const BlockSp<T, B>& A, const solArray<T, B, D>& x, int dim = -1, bool transpose = false) {
    solArray<T, B, D> b(transpose ? A.n() : A.m());
    if (D == 1 && dim < 0) mult_detail::bsmv_device(b, A, x, false, transpose, 0, -1);
    else mult_detail::bsmv_device(b, A, x, false, transpose, 1, dim);
    return b;
}

/// BlockSp<T,B*D> x solArray<T,B,D> (D > 1): the plain block matrix-vector product.
template <typename T, int B, int D>
    requires(D > 1)
solArray<T, B, D> bsmv(const BlockSp<T, B * D>& A, const solArray<T, B, D>& x, bool transpose = false) {
    solArray<T, B, D> b(transpose ? A.n() : A.m());
    mult_detail::bsmv_device(b, A, x, false, transpose, 0, -1);
    return b;
}

/// Void forms: b += A x.
template <typename T, int B, int D = 1>
void bsmv(solArray<T, B, D>& b, const BlockSp<T, B>& A, const solArray<T, B, D>& x, int dim = -1,
          bool transpose = false) {
    if (D == 1 && dim < 0) mult_detail::bsmv_device(b, A, x, true, transpose, 0, -1);
    else mult_detail::bsmv_device(b, A, x, true, transpose, 1, dim);
}

template <typename T, int B, int D>
    requires(D > 1)
void bsmv(solArray<T, B, D>& b, const BlockSp<T, B * D>& A, const solArray<T, B, D>& x, bool transpose = false) {
    mult_detail::bsmv_device(b, A, x, true, transpose, 0, -1);
}

template <typename T, int B, int D = 1>
solArray<T, B, D> operator*(const BlockSp<T, B * D>& A, const solArray<T, B, D>& x) {
    solArray<T, B, D> b(A.m());
    mult_detail::bsmv_device(b, A, x, false, false, 0, -1);
    return b;
}

// ===========================================================================
// bs_id, bs_tr, bsmm (blocksp_multiply.hpp:21-159, 334-363)
// ===========================================================================

namespace mult_detail {

/// Identity of size blocks (multiply.hpp:21-28).
template <typename T, int M>
BlockSp<T, M> blkSpIdent(int size) {
    BlockSp<T, M> I(size);
    for (int im = 0; im < size; ++im) {
        auto& blk = I(im, im);
        for (int i = 0; i < M; ++i) blk(i, i) = T(1);
    }
    return I;
}

/// Transpose with transposed blocks (multiply.hpp:31-44): row j lists the rows i holding a
/// block in column j in ascending i; blocks are conjugate-transposed for complex T, as
/// dense::transpose is (dense_multiply.hpp:207-213).
template <typename T, int M>
BlockSp<T, M> blkSpTranspose(const BlockSp<T, M>& A) {
    BlockSp<T, M> tr(A.n(), A.m());
    for (int im = 0; im < A.m(); ++im)
        for (size_t in = 0; in < A.colSize(im); ++in) {
            const int j = A.colInd(im, (int)in);
            const auto& src = A.get_data(im)[in];
            auto& dst = tr(j, im);
            for (int r = 0; r < M; ++r)
                for (int c = 0; c < M; ++c) {
                    if constexpr (is_complex_v<T>)
                        dst(c, r) = std::conj(src(r, c));
                    else
                        dst(c, r) = src(r, c);
                }
        }
    return tr;
}

/// C = A B or A^T B on the device (multiply.hpp:51-136): C's rows list their blocks in the
/// reference's first-touch order; the product runs in FP64.
template <typename T, int M>
BlockSp<T, M> blkSpMatMat(const BlockSp<T, M>& A, const BlockSp<T, M>& B, bool Atran = false) {
    if (!Atran) assert(A.n() == B.m());
    else assert(A.m() == B.m());
    auto dA = gpu::device_matrix(A);
    auto dB = gpu::device_matrix(B);
    bsp_bsr* h = nullptr;
    const int rc = Atran ? bsp_bsmm_trans_a(gpu::ctx(), dA->h, dB->h, &h) : bsp_bsmm(gpu::ctx(), dA->h, dB->h, &h);
    gpu::check(gpu::ctx(), rc, "bsmm");
    std::unique_ptr<bsp_bsr, int (*)(bsp_bsr*)> guard(h, bsp_bsr_destroy);
    int nr = 0, nc = 0, bs = 0;
    int64_t nz = 0;
    gpu::check(gpu::ctx(), bsp_bsr_info(h, &nr, &nc, &bs, &nz), "bsr_info");
    std::vector<int32_t> rp(nr + 1), ci(std::max<int64_t>(nz, 1));
    std::vector<double> v(std::max<size_t>(gpu::doubles<T>((size_t)nz * M * M), 1));
    if constexpr (is_complex_v<T>)
        gpu::check(gpu::ctx(), bsp_bsr_download_complex(gpu::ctx(), h, rp.data(), ci.data(), v.data()),
                   "bsr_download_complex");
    else
        gpu::check(gpu::ctx(), bsp_bsr_download(gpu::ctx(), h, rp.data(), ci.data(), v.data()), "bsr_download");
    BlockSp<T, M> C(nr, nc);
    for (int i = 0; i < nr; ++i) {
        C.reserve(i, rp[i + 1] - rp[i]);
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            auto& blk = C(i, ci[p]);
            for (int e = 0; e < M * M; ++e) {
                const size_t q = (size_t)p * M * M + e;
                if constexpr (is_complex_v<T>)
                    blk.data()[e] = T((ScalarType_t<T>)v[2 * q], (ScalarType_t<T>)v[2 * q + 1]);
                else
                    blk.data()[e] = static_cast<T>(v[q]);
            }
        }
    }
    return C;
}

/// D = A B C or A^T B C (multiply.hpp:138-147).
template <typename T, int M>
BlockSp<T, M> blkSpMatMat(const BlockSp<T, M>& A, const BlockSp<T, M>& B, const BlockSp<T, M>& C,
                          bool Atran = false) {
    return blkSpMatMat<T, M>(A, blkSpMatMat<T, M>(B, C), Atran);
}

/// E = A B C D or A^T B C D (multiply.hpp:149-159).
template <typename T, int M>
BlockSp<T, M> blkSpMatMat(const BlockSp<T, M>& A, const BlockSp<T, M>& B, const BlockSp<T, M>& C,
                          const BlockSp<T, M>& D, bool Atran = false) {
    return blkSpMatMat<T, M>(A, blkSpMatMat<T, M>(B, C, D), Atran);
}

}  // namespace mult_detail

template <typename T, int M>
BlockSp<T, M> bs_id(int size) { return mult_detail::blkSpIdent<T, M>(size); }
template <typename T, int M>
BlockSp<T, M> bs_tr(const BlockSp<T, M>& A) { return mult_detail::blkSpTranspose<T, M>(A); }
template <typename T, int M>
BlockSp<T, M> bsmm(const BlockSp<T, M>& A, const BlockSp<T, M>& B, bool Atran = false) {
    return mult_detail::blkSpMatMat<T, M>(A, B, Atran);
}
template <typename T, int M>
BlockSp<T, M> bsmm(const BlockSp<T, M>& A, const BlockSp<T, M>& B, const BlockSp<T, M>& C, bool Atran = false) {
    return mult_detail::blkSpMatMat<T, M>(A, B, C, Atran);
}
template <typename T, int M>
BlockSp<T, M> bsmm(const BlockSp<T, M>& A, const BlockSp<T, M>& B, const BlockSp<T, M>& C, const BlockSp<T, M>& D,
                   bool Atran = false) {
    return mult_detail::blkSpMatMat<T, M>(A, B, C, D, Atran);
}

// ===========================================================================
// preconditioners (blocksp_preconditioners.hpp)
// ===========================================================================

namespace dense {
/// Diagonal-block solver tags (dense_solvers.hpp:47-190): LUFactor = LU with partial
/// pivoting, CholeskyFactor = Cholesky (SPD blocks).
template <typename T, int N> struct LUFactor { static constexpr int kind = 0; };
template <typename T, int N> struct CholeskyFactor { static constexpr int kind = 1; };

/// Eigenvalue set of an Arnoldi Hessenberg (dense_solvers.hpp:909-1021): real and
/// imaginary parts, magnitudes and the same summary statistics and printouts.
template <typename T>
struct Eigenvalues_Arr {
    std::vector<double> WR, WI, WM;
    bool harmonic_Ritz = false;
    Eigenvalues_Arr() {}
    Eigenvalues_Arr(std::vector<double> re, std::vector<double> im, bool harmonic)
        : WR(std::move(re)), WI(std::move(im)), harmonic_Ritz(harmonic) {
        WM.resize(WR.size());
        for (size_t i = 0; i < WR.size(); ++i) WM[i] = std::sqrt(WR[i] * WR[i] + WI[i] * WI[i]);
    }
    const std::vector<double>& real() const { return WR; }
    const std::vector<double>& imag() const { return WI; }
    const std::vector<double>& mag() const { return WM; }
    static double vmin(const std::vector<double>& v) { return *std::min_element(v.begin(), v.end()); }
    static double vmax(const std::vector<double>& v) { return *std::max_element(v.begin(), v.end()); }
    double min_real() const { return vmin(WR); }
    double max_real() const { return vmax(WR); }
    double min_imag() const { return vmin(WI); }
    double max_imag() const { return vmax(WI); }
    double min_mag() const { return vmin(WM); }
    double max_mag() const { return vmax(WM); }
    int arg_min_mag() const { return (int)(std::min_element(WM.begin(), WM.end()) - WM.begin()); }
    int arg_max_mag() const { return (int)(std::max_element(WM.begin(), WM.end()) - WM.begin()); }
    double cond_number() const { return max_mag() / min_mag(); }
    void printEigVals() const {
        auto pr = [](const std::vector<double>& v) {
            for (double e : v) std::cout << e << ' ';
            std::cout << std::endl;
        };
        if (harmonic_Ritz) std::cout << "Harmonic Ritz values" << std::endl;
        std::cout << "Eigenvalues - real component" << std::endl;
        pr(WR);
        std::cout << "Eigenvalues - imaginary component" << std::endl;
        pr(WI);
        std::cout << "Eigenvalues - magnitude" << std::endl;
        pr(WM);
    }
    void printEigInfo() const {
        if (WR.empty()) return;
        if (harmonic_Ritz) std::cout << "Harmonic Ritz info" << std::endl;
        else std::cout << "Eigenvalue info" << std::endl;
        std::cout << "min/max real eig " << min_real() << ' ' << max_real() << std::endl;
        std::cout << "min/max img  eig " << min_imag() << ' ' << max_imag() << std::endl;
        std::cout << "smallest magn eig " << min_mag() << std::endl;
        std::cout << "largest magn eig  " << max_mag() << std::endl;
        std::cout << "ratio/cond " << cond_number() << std::endl;
    }
};
}  // namespace dense

namespace preConditioners {

/// Base of the device-backed preconditioners: operator() applies M^{-1} in place.
template <typename T, int B, int D>
class DeviceBase {
protected:
    std::shared_ptr<gpu::DPrecond> pre_;

public:
    using is_device_precond = void;
    bsp_precond* handle() const { return pre_ ? pre_->h : nullptr; }
    void operator()(solArray<T, B, D>& r) const {
        if (!pre_) return;
        gpu::DVec dr(gpu::doubles<T>(r.size_flat()));
        dr.upload(r.data());
        gpu::check(gpu::ctx(), bsp_precond_apply(gpu::ctx(), pre_->h, dr.get()), "precond_apply");
        dr.download(r.data());
    }
};

/// NoPre (preconditioners.hpp:66-71).
template <typename T, int B, int D = 1>
struct NoPre {
    NoPre() {}
    inline void operator()(solArray<T, B, D>&) const {}
};

/// Diag<Inv> (preconditioners.hpp:75-111).
template <typename Inv, typename T, int B, int D = 1>
struct Diag : DeviceBase<T, B, D> {
    int len = 0;
    const BlockSp<T, B * D>* A_Ptr = nullptr;
    Diag() {}
    explicit Diag(const BlockSp<T, B * D>& A) : len(A.m()), A_Ptr(&A) {
        auto p = std::make_shared<gpu::DPrecond>();
        p->A = gpu::device_matrix(A);
        gpu::check(gpu::ctx(), bsp_precond_create_diag(gpu::ctx(), p->A->h, Inv::kind, &p->h), "Diag");
        this->pre_ = p;
    }
};

/// SOR<Inv>(A, w, nu, type) (preconditioners.hpp:116-240): from x0 = 0, nu sweeps per
/// direction; type < 0 backward, 0 symmetric, > 0 forward.
template <typename Inv, typename T, int B, int D = 1>
struct SOR : DeviceBase<T, B, D> {
    const BlockSp<T, B * D>* A_Ptr = nullptr;
    ScalarType_t<T> m_w = 0;
    int m_nu = 0;
    int m_type = 0;
    SOR() {}
    SOR(const BlockSp<T, B * D>& A, ScalarType_t<T> w, int nu, int type = 0) : A_Ptr(&A), m_w(w), m_nu(nu), m_type(type) {
        rebuild();
    }
    void rebuild() {
        auto p = std::make_shared<gpu::DPrecond>();
        p->A = gpu::device_matrix(*A_Ptr);
        gpu::check(gpu::ctx(), bsp_precond_create_sor(gpu::ctx(), p->A->h, (double)m_w, m_nu, m_type, Inv::kind, &p->h),
                   "SOR");
        this->pre_ = p;
    }
    // The reference's change_w takes an int (preconditioners.hpp:236); kept as is.
    inline void change_w(int w_new) { m_w = w_new; rebuild(); }
    inline void change_nu(int nu_new) { m_nu = nu_new; rebuild(); }
    inline void change_type(int type_new) { m_type = type_new; rebuild(); }
};

/// ILU0(A) (preconditioners.hpp:246-357).
template <typename T, int B, int D = 1>
struct ILU0 : DeviceBase<T, B, D> {
    int n = 0;
    ILU0() {}
    explicit ILU0(const BlockSp<T, B * D>& A) : n(A.n()) {
        auto p = std::make_shared<gpu::DPrecond>();
        p->A = gpu::device_matrix(A);
        gpu::check(gpu::ctx(), bsp_precond_create_ilu0(gpu::ctx(), p->A->h, &p->h), "ILU0");
        this->pre_ = p;
    }
};

/// IChol0(A) (preconditioners.hpp:364-538), A symmetric positive definite.
template <typename T, int B, int D = 1>
struct IChol0 : DeviceBase<T, B, D> {
    int n = 0;
    IChol0() {}
    explicit IChol0(const BlockSp<T, B * D>& A) : n(A.n()) {
        auto p = std::make_shared<gpu::DPrecond>();
        p->A = gpu::device_matrix(A);
        gpu::check(gpu::ctx(), bsp_precond_create_ichol0(gpu::ctx(), p->A->h, &p->h), "IChol0");
        this->pre_ = p;
    }
};

}  // namespace preConditioners

// ===========================================================================
// Krylov solvers (blocksp_solvers.hpp)
// ===========================================================================

namespace solver_detail {

template <typename P>
concept DevicePrecond = requires(const P& p) { p.handle(); };

template <typename T, int B, int D, typename PreCond>
struct HostPrecondThunk {
    const PreCond* pc;
    // r: len doubles, (re, im) pairs for complex T
    static void call(double* r, int64_t len, void* user) {
        using S = ScalarType_t<T>;
        constexpr int64_t per = is_complex_v<T> ? 2 : 1;
        auto* self = static_cast<HostPrecondThunk*>(user);
        solArray<T, B, D> sa((int)(len / per / (B * D)));
        S* s = reinterpret_cast<S*>(sa.data());
        for (int64_t i = 0; i < len; ++i) s[i] = static_cast<S>(r[i]);
        (*self->pc)(sa);
        for (int64_t i = 0; i < len; ++i) r[i] = static_cast<double>(s[i]);
    }
};

// Runs a device solver with the preconditioner handle `PreCond` provides, wrapping
// host-only preconditioners (lambdas, NoPre, user structs) in a host callback.
template <typename T, int B, int D, typename PreCond, typename Fn>
solArray<T, B, D> run(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, const PreCond& preCond, Fn&& fn) {
    auto dA = gpu::device_matrix(A);
    gpu::DPrecond wrapper;
    bsp_precond* h = nullptr;
    HostPrecondThunk<T, B, D, PreCond> thunk{&preCond};
    if constexpr (DevicePrecond<PreCond>) {
        h = preCond.handle();
    }
    if (!h) {
        if constexpr (std::is_same_v<PreCond, preConditioners::NoPre<T, B, D>>) {
            gpu::check(gpu::ctx(), bsp_precond_create_none(gpu::ctx(), dA->h, &wrapper.h), "NoPre");
        } else {
            gpu::check(gpu::ctx(),
                       bsp_precond_create_host(gpu::ctx(), dA->h, &HostPrecondThunk<T, B, D, PreCond>::call, &thunk,
                                               &wrapper.h),
                       "host preconditioner");
        }
        h = wrapper.h;
    }
    gpu::DVec db(gpu::doubles<T>(b.size_flat())), dx(gpu::doubles<T>(b.size_flat()));
    db.upload(b.data());
    fn(dA->h, h, db.get(), dx.get());
    solArray<T, B, D> x(b.len());
    dx.download(x.data());
    return x;
}

}  // namespace solver_detail

/// Preconditioned conjugate gradient (blocksp_solvers.hpp:51-105).
template <typename T, int B, int D, typename PreCond>
solArray<T, B, D> pCG(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, T tol, const PreCond& preCond,
                      int maxIters = 1000, bool printIters = false) {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>);
    bsp_stats st{};
    auto x = solver_detail::run<T, B, D>(A, b, preCond, [&](bsp_bsr* a, bsp_precond* m, double* db, double* dx) {
        int rc = bsp_pcg(gpu::ctx(), a, m, db, dx, (double)tol, maxIters, BSP_TOL_REFERENCE, &st);
        gpu::check(gpu::ctx(), rc, "pCG");
    });
    if (printIters) {
        std::cout << "pCG iters " << st.iters << std::endl;
        if (st.iters > maxIters) std::cout << "maximum pCG iterations reached" << std::endl;
    }
    return x;
}

/// Left-preconditioned GMRES with deflated restart (blocksp_solvers.hpp:156-295).
template <typename T, int B, int D, typename PreCond>
solArray<T, B, D> pGMRES(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, ScalarType_t<T> tol,
                         const PreCond& preCond, int maxIters = 500, int restart = -1, int k = -1,
                         bool printIters = false, bool printEigInfo = false) {
    bsp_stats st{};
    auto x = solver_detail::run<T, B, D>(A, b, preCond, [&](bsp_bsr* a, bsp_precond* m, double* db, double* dx) {
        int rc = bsp_pgmres(gpu::ctx(), a, m, db, dx, (double)tol, maxIters, restart, k, BSP_TOL_REFERENCE, &st);
        gpu::check(gpu::ctx(), rc, "pGMRES");
    });
    if (printIters) {
        std::cout << "pGMRES iters " << st.iters << std::endl;
        if (st.iters > maxIters) std::cout << "maximum pGMRES iterations reached" << std::endl;
    }
    if (printEigInfo) {
        // eigenvalues of the final Hessenberg's square part; harmonic Ritz values if restarted
        int cnt = 0, harm = 0;
        gpu::check(gpu::ctx(), bsp_pgmres_ritz(gpu::ctx(), nullptr, nullptr, 0, &cnt, &harm), "pgmres_ritz");
        std::vector<double> re(cnt), im(cnt);
        gpu::check(gpu::ctx(), bsp_pgmres_ritz(gpu::ctx(), re.data(), im.data(), cnt, &cnt, &harm), "pgmres_ritz");
        dense::Eigenvalues_Arr<T>(std::move(re), std::move(im), harm != 0).printEigInfo();
    }
    return x;
}

/// Arnoldi eigenvalue estimates of M^{-1} A (blocksp_solvers.hpp:298-333): `iters` Arnoldi
/// steps from q_0 = e_0 on the device; (harmonic) Ritz values of the Hessenberg.
template <typename T, int M, typename PreCond = preConditioners::NoPre<T, M, 1>>
dense::Eigenvalues_Arr<T> arnoldi_eigenvalues(const BlockSp<T, M>& A, int iters = 30,
                                              const PreCond* preCond = nullptr, bool printEigInfo = false,
                                              bool printEigVals = false, bool harmonic_Ritz = false) {
    auto dA = gpu::device_matrix(A);
    gpu::DPrecond wrapper;
    bsp_precond* h = nullptr;
    solver_detail::HostPrecondThunk<T, M, 1, PreCond> thunk{preCond};
    if (preCond) {
        if constexpr (solver_detail::DevicePrecond<PreCond>) h = preCond->handle();
        if (!h && !std::is_same_v<PreCond, preConditioners::NoPre<T, M, 1>>) {
            gpu::check(gpu::ctx(),
                       bsp_precond_create_host(gpu::ctx(), dA->h, &solver_detail::HostPrecondThunk<T, M, 1, PreCond>::call,
                                               &thunk, &wrapper.h),
                       "host preconditioner");
            h = wrapper.h;
        }
    }
    std::vector<double> re(iters), im(iters);
    int cnt = 0;
    gpu::check(gpu::ctx(),
               bsp_arnoldi_eigenvalues(gpu::ctx(), dA->h, h, iters, harmonic_Ritz ? 1 : 0, re.data(), im.data(), iters,
                                       &cnt),
               "arnoldi_eigenvalues");
    re.resize(cnt);
    im.resize(cnt);
    dense::Eigenvalues_Arr<T> eig(std::move(re), std::move(im), harmonic_Ritz);
    if (printEigInfo) eig.printEigInfo();
    if (printEigVals) eig.printEigVals();
    return eig;
}

/// pGMRES with block SOR preconditioner / LU block inversion (solvers.hpp:344-391).
template <typename T, int B, int D = 1>
struct SOR_pGMRES {
    preConditioners::SOR<dense::LUFactor<T, B * D>, T, B, D> SOR_pre;
    ScalarType_t<T> m_w;
    int m_nu, m_SORtype;
    ScalarType_t<T> m_tol;
    int maxIters, restart, m_k;
    bool printIters, printEigInfo;
    SOR_pGMRES(const BlockSp<T, B * D>& A, ScalarType_t<T> w = 1.6, int nu = 3, int SORtype = 0,
               ScalarType_t<T> tol = 1.0e-8, int maxIters_ = 5000, int restart_ = -1, int k = -1, bool printIters_ = false,
               bool printEigInfo_ = false)
        : m_w(w), m_nu(nu), m_SORtype(SORtype), m_tol(tol), maxIters(maxIters_), restart(restart_), m_k(k),
          printIters(printIters_), printEigInfo(printEigInfo_) {
        SOR_pre = preConditioners::SOR<dense::LUFactor<T, B * D>, T, B, D>(A, m_w, m_nu, m_SORtype);
    }
    solArray<T, B, D> solve(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b) const {
        return pGMRES<T, B, D>(A, b, m_tol, SOR_pre, maxIters, restart, m_k, printIters, printEigInfo);
    }
    auto& access_SOR_pre() { return SOR_pre; }
    void change_SOR_pre(const preConditioners::SOR<dense::LUFactor<T, B * D>, T, B, D>& s) { SOR_pre = s; }
    // As in the reference (solvers.hpp:382), changing w here does not rebuild SOR_pre.
    void change_w(ScalarType_t<T> w) { m_w = w; }
    void change_nu(int nu) { m_nu = nu; }
    void change_SORtype(int t) { m_SORtype = t; }
    void change_tol(ScalarType_t<T> tol) { m_tol = tol; }
    void change_maxIters(int it) { maxIters = it; }
    void change_restart(int re) { restart = re; }
    void change_k(int k) { m_k = k; }
    void change_printIters(bool p) { printIters = p; }
    void change_printEigInfo(bool p) { printEigInfo = p; }
};

/// pCG with block SOR preconditioner / Cholesky block inversion (solvers.hpp:395-431).
template <typename T, int B, int D = 1>
struct SOR_pCG {
    preConditioners::SOR<dense::CholeskyFactor<T, B * D>, T, B, D> SOR_pre;
    ScalarType_t<T> m_w;
    int m_nu;
    ScalarType_t<T> m_tol;
    int maxIters;
    int printIters;
    SOR_pCG(const BlockSp<T, B * D>& A, ScalarType_t<T> w = 1.6, int nu = 3, ScalarType_t<T> tol = 1.0e-8,
            int maxIters_ = 5000, bool printIters_ = false)
        : m_w(w), m_nu(nu), m_tol(tol), maxIters(maxIters_), printIters(printIters_) {
        static_assert(!is_complex_v<T>);
        SOR_pre = preConditioners::SOR<dense::CholeskyFactor<T, B * D>, T, B, D>(A, m_w, m_nu);
    }
    solArray<T, B, D> solve(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b) {
        return pCG<T, B, D>(A, b, m_tol, SOR_pre, maxIters, printIters);
    }
    auto& access_SOR_pre() { return SOR_pre; }
    void change_SOR_pre(const preConditioners::SOR<dense::CholeskyFactor<T, B * D>, T, B, D>& s) { SOR_pre = s; }
    void change_w(ScalarType_t<T> w) { m_w = w; }
    void change_nu(int nu) { m_nu = nu; }
    void change_tol(ScalarType_t<T> tol) { m_tol = tol; }
    void change_maxIters(int it) { maxIters = it; }
    void change_printIters(bool p) { printIters = p; }
};

/// pGMRES with block ILU0 (solvers.hpp:434-467).
template <typename T, int B, int D = 1>
struct ILU0_pGMRES {
    ScalarType_t<T> m_tol;
    int maxIters, restart, m_k;
    bool printIters, printEigInfo;
    preConditioners::ILU0<T, B, D> ILU0_pre;
    ILU0_pGMRES(const BlockSp<T, B * D>& A, ScalarType_t<T> tol = 1.0e-8, int maxIters_ = 5000, int restart_ = -1,
                int k = -1, bool printIters_ = false, bool printEigInfo_ = false)
        : m_tol(tol), maxIters(maxIters_), restart(restart_), m_k(k), printIters(printIters_),
          printEigInfo(printEigInfo_), ILU0_pre(A) {}
    solArray<T, B, D> solve(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b) const {
        return pGMRES<T, B, D>(A, b, m_tol, ILU0_pre, maxIters, restart, m_k, printIters, printEigInfo);
    }
    auto& access_ILU0_pre() { return ILU0_pre; }
    void change_tol(ScalarType_t<T> tol) { m_tol = tol; }
    void change_maxIters(int it) { maxIters = it; }
    void change_restart(int re) { restart = re; }
    void change_k(int k) { m_k = k; }
    void change_printIters(bool p) { printIters = p; }
    void change_printEigInfo(bool p) { printEigInfo = p; }
};

/// pCG with block IChol0 (solvers.hpp:471-498).
template <typename T, int B, int D = 1>
struct IChol0_pCG {
    ScalarType_t<T> m_tol;
    int maxIters;
    int printIters;
    preConditioners::IChol0<T, B, D> IChol0_pre;
    IChol0_pCG(const BlockSp<T, B * D>& A, ScalarType_t<T> tol = 1.0e-8, int maxIters_ = 5000, bool printIters_ = false)
        : m_tol(tol), maxIters(maxIters_), printIters(printIters_), IChol0_pre(A) {
        static_assert(!is_complex_v<T>);
    }
    solArray<T, B, D> solve(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b) {
        return pCG<T, B, D>(A, b, m_tol, IChol0_pre, maxIters, printIters);
    }
    auto& access_IChol0_pre() { return IChol0_pre; }
    void change_tol(ScalarType_t<T> tol) { m_tol = tol; }
    void change_maxIters(int it) { maxIters = it; }
    void change_printIters(bool p) { printIters = p; }
};

/// One-call solvers with the reference's defaults (solvers.hpp:508-649).
namespace ez {

template <typename T> constexpr ScalarType_t<T> default_tol() {
    return (std::is_same_v<T, float> || std::is_same_v<T, std::complex<float>>) ? 1.0e-5 : 1.0e-8;
}

template <typename T, int B, int D = 1>
solArray<T, B, D> sor_pGMRES(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false,
                             bool printEigInfo = false) {
    return SOR_pGMRES<T, B, D>(A, 1.6, 3, 0, default_tol<T>(), 5000, -1, -1, printIters, printEigInfo).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> sor_pGMRES_robust(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false,
                                    bool printEigInfo = false) {
    return SOR_pGMRES<T, B, D>(A, 0.6, 3, 0, default_tol<T>(), 5000, -1, -1, printIters, printEigInfo).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> sor_pGMRES_dr(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false,
                                bool printEigInfo = false) {
    return SOR_pGMRES<T, B, D>(A, 1.6, 3, 0, default_tol<T>(), 5000, 30, 10, printIters, printEigInfo).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> ilu0_pGMRES(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false,
                              bool printEigInfo = false) {
    return ILU0_pGMRES<T, B, D>(A, default_tol<T>(), 5000, -1, -1, printIters, printEigInfo).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> ilu0_pGMRES_dr(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false,
                                 bool printEigInfo = false) {
    return ILU0_pGMRES<T, B, D>(A, default_tol<T>(), 5000, 30, 10, printIters, printEigInfo).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> sor_pCG(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false) {
    return SOR_pCG<T, B, D>(A, 1.6, 3, default_tol<T>(), 5000, printIters).solve(A, b);
}
template <typename T, int B, int D = 1>
solArray<T, B, D> ichol0_pCG(const BlockSp<T, B * D>& A, const solArray<T, B, D>& b, bool printIters = false) {
    return IChol0_pCG<T, B, D>(A, default_tol<T>(), 5000, printIters).solve(A, b);
}

}  // namespace ez

// ===========================================================================
// Matrix-Market converters (blocksp_tests.hpp:578-699, namespace tests)
// ===========================================================================

namespace tests {

/// GotoLine (blocksp_tests.hpp:582-590): rewinds and skips num - 1 lines.
inline std::ifstream& GotoLine(std::ifstream& file, int num) {
    file.seekg(0, std::ios::beg);
    for (int i = 0; i < num - 1; ++i) file.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    return file;
}

/// mtxMatrix_to_BlockSp (blocksp_tests.hpp:595-649): line start_line holds "M N nnz", the
/// nnz lines after it "i j value" (complex T: "i j re im"), 1-based.  Every entry goes
/// through setFromGlobal, so a block row lists its blocks in first-touch order of the
/// file; with symmetric the mirrored entry (j, i) -- conjugated for complex T -- is written
/// right after (i, j) when i != j; a repeated entry keeps the last value.  As in the
/// reference, base_zero is accepted and not used: indices are always read as 1-based.
/// Unlike the reference (an assert, or garbage from an unopened stream), a file that cannot
/// be read or a size that B does not divide throws BlockSp::Error.
template <typename T, int B = 1>
BlockSp<T, B> mtxMatrix_to_BlockSp(const std::string& mat_loc, unsigned int start_line = 0, bool symmetric = false,
                                   bool base_zero = false, bool printMatSize = false) {
    (void)base_zero;
    std::ifstream f(mat_loc, std::ios::in);
    if (!f) throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: cannot open " + mat_loc);
    GotoLine(f, (int)start_line);
    std::string size;
    std::getline(f, size);
    std::stringstream ss(size);
    int M = -1, N = -1, nnz = -1;
    ss >> M >> N >> nnz;
    if (!ss || M < 0 || N < 0 || nnz < 0)
        throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: no \"M N nnz\" line in " + mat_loc);
    if (printMatSize) {
        std::cout << M << ' ' << N << ' ' << nnz;
        if (symmetric) std::cout << " Symmetric";
        std::cout << std::endl;
    }
    if (M % B != 0 || N % B != 0)
        throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: block size " + std::to_string(B) +
                                       " does not divide " + std::to_string(M) + " x " + std::to_string(N));
    BlockSp<T, B> A(M / B, N / B);
    for (int e = 0; e < nnz; ++e) {
        std::string entry;
        if (!std::getline(f, entry)) throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: file ends early: " + mat_loc);
        std::stringstream es(entry);
        int i = 0, j = 0;
        es >> i >> j;
        --i;
        --j;
        if (!es || i < 0 || j < 0 || i >= M || j >= N)
            throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: bad entry \"" + entry + "\"");
        T val;
        if constexpr (is_complex_v<T>) {
            ScalarType_t<T> re = 0, im = 0;
            es >> re >> im;
            val = T(re, im);
        } else {
            es >> val;
        }
        if (!es) throw Error(BSP_E_INVALID, "mtxMatrix_to_BlockSp: bad value \"" + entry + "\"");
        util::setFromGlobal(A, i, j, val);
        if (symmetric && i != j) {
            if constexpr (is_complex_v<T>)
                util::setFromGlobal(A, j, i, std::conj(val));
            else
                util::setFromGlobal(A, j, i, val);
        }
    }
    return A;
}

/// mtxArray_to_solArray (blocksp_tests.hpp:652-699): line start_line holds "M 1", the M lines
/// after it one value each (complex T: "re im"); value i goes to solArray element
/// (i / (B D), dim, ib) with i % (B D) = dim B + ib, the layout setFromGlobal uses
/// (blocksp_utility.hpp:630-639).  The reference computes the in-element index as i % B
/// (blocksp_tests.hpp:679), which for D > 1 writes every value to dim 0; for D = 1 the two
/// agree.  Errors throw BlockSp::Error as in mtxMatrix_to_BlockSp.
template <typename T, int B = 1, int D = 1>
solArray<T, B, D> mtxArray_to_solArray(const std::string& vec_loc, unsigned int start_line = 0,
                                       bool printVecSize = false) {
    std::ifstream f(vec_loc, std::ios::in);
    if (!f) throw Error(BSP_E_INVALID, "mtxArray_to_solArray: cannot open " + vec_loc);
    GotoLine(f, (int)start_line);
    std::string size;
    std::getline(f, size);
    std::stringstream ss(size);
    int M = -1, N = -1;
    ss >> M >> N;
    if (!ss || M < 0) throw Error(BSP_E_INVALID, "mtxArray_to_solArray: no \"M N\" line in " + vec_loc);
    if (printVecSize) std::cout << M << ' ' << N << std::endl;
    constexpr int BD = B * D;
    if (M % BD != 0 || N != 1)
        throw Error(BSP_E_INVALID, "mtxArray_to_solArray: need one column and B*D = " + std::to_string(BD) +
                                       " dividing " + std::to_string(M));
    solArray<T, B, D> x(M / BD);
    for (int i = 0; i < M; ++i) {
        std::string entry;
        if (!std::getline(f, entry)) throw Error(BSP_E_INVALID, "mtxArray_to_solArray: file ends early: " + vec_loc);
        std::stringstream es(entry);
        T val;
        if constexpr (is_complex_v<T>) {
            long double re = 0, im = 0;
            es >> re >> im;
            val = T((ScalarType_t<T>)re, (ScalarType_t<T>)im);
        } else {
            es >> val;
        }
        if (!es) throw Error(BSP_E_INVALID, "mtxArray_to_solArray: bad value \"" + entry + "\"");
        util::setFromGlobal(x, i, val);
    }
    return x;
}

}  // namespace tests

}  // namespace BlockSp
