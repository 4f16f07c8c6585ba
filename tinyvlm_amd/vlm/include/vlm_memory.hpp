// vlm_memory.hpp -- the tensor data model every Backend call passes (row a16 of the hot path).
//
// Same contract as the reference (vlm/include/vlm_memory.hpp:22-372): column-major strided views
// (stride[0] == 1 for a fresh tensor, stride[i] = stride[i-1] * shape[i-1]), non-owning
// TensorView with slice / reshape / to / fill, an owning Tensor allocated through a Memory
// manager, and MultiTensor holding one tensor per surface.  Negative indices count from the end;
// Range{first, last} with last = -1 meaning "to the end" (All = {0, -1}).
#pragma once

#include <array>
#include <cassert>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "vlm_types.hpp"

namespace vlm {

enum class Location { Device, Host };

class Memory {
   public:
    explicit Memory(bool unified) : m_unified(unified) {}
    virtual ~Memory() = default;
    virtual void* alloc(Location location, i64 size_bytes) const = 0;
    virtual void free(Location location, void* ptr) const = 0;
    virtual void copy(Location dst_loc, void* dst, i64 dst_stride, Location src_loc, const void* src, i64 src_stride,
                      i64 elem_size, i64 size) const = 0;
    virtual void fill(Location loc, float* ptr, i64 stride, float value, i64 size) const = 0;
    virtual void fill(Location loc, double* ptr, i64 stride, double value, i64 size) const = 0;
    bool is_unified() const { return m_unified; }

   private:
    const bool m_unified;
};

using Range = std::array<i64, 2>;
constexpr Range All{0, -1};

template <typename T, i64 Dim, Location L>
class TensorView {
   public:
    using DimArray = std::array<i64, Dim>;

    TensorView() = default;
    TensorView(Memory* memory, const DimArray& shape, const DimArray& stride, T* ptr)
        : m_memory(memory), m_shape(shape), m_stride(stride), m_ptr(ptr) {}

    i64 size() const {
        i64 s = 1;
        for (i64 d = 0; d < Dim; d++) s *= m_shape[d];
        return s;
    }
    i64 size_bytes() const { return size() * (i64)sizeof(T); }
    const DimArray& shape() const { return m_shape; }
    const DimArray& stride() const { return m_stride; }
    i64 shape(i64 d) const { return m_shape[d]; }
    i64 stride(i64 d) const { return m_stride[d]; }
    T* ptr() const { return m_ptr; }
    Memory* memory() const { return m_memory; }
    static constexpr Location location() { return L; }
    static constexpr i64 dim() { return Dim; }

    i64 offset(const DimArray& idx) const {
        i64 o = 0;
        for (i64 d = 0; d < Dim; d++) o += (idx[d] < 0 ? idx[d] + m_shape[d] : idx[d]) * m_stride[d];
        return o;
    }
    template <typename... Idx>
    T& operator()(Idx... idx) const {
        static_assert(sizeof...(Idx) == Dim, "index count must equal the dimension");
        return m_ptr[offset(DimArray{static_cast<i64>(idx)...})];
    }
    T& operator[](i64 i) const { return m_ptr[i]; }

    // slice: each argument is an integer (dimension dropped) or a Range (dimension kept)
    template <typename... Args>
    auto slice(Args... args) const {
        static_assert(sizeof...(Args) == Dim, "one argument per dimension");
        constexpr i64 M = (0 + ... + (std::is_same_v<std::decay_t<Args>, Range> ? 1 : 0));
        std::array<i64, M> nshape{}, nstride{};
        T* p = m_ptr;
        i64 d = 0, k = 0;
        auto one = [&](auto a) {
            if constexpr (std::is_same_v<std::decay_t<decltype(a)>, Range>) {
                const i64 first = a[0] < 0 ? m_shape[d] + a[0] : a[0];
                const i64 last = a[1] < 0 ? m_shape[d] + a[1] + 1 : a[1];
                if (first < 0 || last > m_shape[d] || last <= first) throw std::out_of_range("slice range");
                p += first * m_stride[d];
                nshape[k] = last - first;
                nstride[k] = m_stride[d];
                k++;
            } else {
                const i64 i = a < 0 ? m_shape[d] + a : a;
                if (i < 0 || i >= m_shape[d]) throw std::out_of_range("slice index");
                p += i * m_stride[d];
            }
            d++;
        };
        (one(args), ...);
        return TensorView<T, M, L>(m_memory, nshape, nstride, p);
    }

    // reshape without copy: consecutive old dimensions are regrouped into new ones; every group of
    // old dimensions that is split or merged must be contiguous (stride[k+1] == stride[k]*shape[k]).
    template <typename... Args>
    auto reshape(Args... args) const {
        constexpr i64 D = sizeof...(Args);
        std::array<i64, D> ns{static_cast<i64>(args)...}, nst{};
        i64 prod = 1;
        for (i64 x : ns) prod *= x;
        if (prod != size()) throw std::invalid_argument("reshape: size mismatch");
        // old dimensions of extent 1 carry no layout information
        std::array<i64, Dim> os{}, ost{};
        i64 on = 0;
        for (i64 d = 0; d < Dim; d++)
            if (m_shape[d] != 1) {
                os[on] = m_shape[d];
                ost[on] = m_stride[d];
                on++;
            }
        i64 oi = 0, oj = 1, ni = 0, nj = 1;
        while (ni < D && oi < on) {
            i64 np = ns[ni], op = os[oi];
            while (np != op) {  // grow the smaller side until both groups hold the same extent
                if (np < op) np *= ns[nj++];
                else op *= os[oj++];
            }
            for (i64 k = oi; k + 1 < oj; k++)
                if (ost[k + 1] != ost[k] * os[k]) throw std::invalid_argument("reshape: non-contiguous group");
            nst[ni] = ost[oi];
            for (i64 k = ni + 1; k < nj; k++) nst[k] = nst[k - 1] * ns[k - 1];
            ni = nj++;
            oi = oj++;
        }
        for (; ni < D; ni++) nst[ni] = ni > 0 ? nst[ni - 1] * ns[ni - 1] : (Dim > 0 ? m_stride[0] : 1);
        return TensorView<T, D, L>(m_memory, ns, nst, m_ptr);
    }

    // element-wise copy into dst (same shape), dimension 0 as strided runs
    template <Location ML>
    void to(const TensorView<T, Dim, ML>& dst) const {
        if (dst.shape() != m_shape) throw std::invalid_argument("to: shape mismatch");
        DimArray idx{};
        copy_rec(dst, idx, Dim - 1);
    }

    void fill(T value) const {
        DimArray idx{};
        fill_rec(value, idx, Dim - 1);
    }

    Memory* m_memory = nullptr;

   private:
    template <Location ML>
    void copy_rec(const TensorView<T, Dim, ML>& dst, DimArray& idx, i64 d) const {
        if (d == 0) {
            m_memory->copy(ML, dst.ptr() + dst.offset(idx), dst.stride(0), L, m_ptr + offset(idx), m_stride[0],
                           (i64)sizeof(T), m_shape[0]);
            return;
        }
        for (i64 i = 0; i < m_shape[d]; i++) {
            idx[d] = i;
            copy_rec(dst, idx, d - 1);
        }
        idx[d] = 0;
    }
    void fill_rec(T value, DimArray& idx, i64 d) const {
        if (d == 0) {
            m_memory->fill(L, m_ptr + offset(idx), m_stride[0], value, m_shape[0]);
            return;
        }
        for (i64 i = 0; i < m_shape[d]; i++) {
            idx[d] = i;
            fill_rec(value, idx, d - 1);
        }
        idx[d] = 0;
    }

    DimArray m_shape{};
    DimArray m_stride{};
    T* m_ptr = nullptr;
};

template <typename T, int Dim, Location L>
class Tensor {
    using View = TensorView<T, Dim, L>;
    View m_view;

   public:
    explicit Tensor(Memory* memory) : m_view(memory, {}, {}, nullptr) {}
    ~Tensor() {
        if (m_view.ptr()) m_view.m_memory->free(L, m_view.ptr());
    }
    Tensor(const Tensor&) = delete;
    Tensor& operator=(const Tensor&) = delete;
    Tensor(Tensor&& o) noexcept : m_view(o.m_view) { o.m_view = View(m_view.m_memory, {}, {}, nullptr); }

    void init(const std::array<i64, Dim>& shape) {
        if (m_view.ptr() && shape == m_view.shape()) return;
        if (m_view.ptr()) m_view.m_memory->free(L, m_view.ptr());
        std::array<i64, Dim> stride{};
        i64 size = shape[0];
        stride[0] = 1;
        for (int i = 1; i < Dim; i++) {
            size *= shape[i];
            stride[i] = stride[i - 1] * shape[i - 1];
        }
        T* p = static_cast<T*>(m_view.m_memory->alloc(L, size * (i64)sizeof(T)));
        m_view = View(m_view.m_memory, shape, stride, p);
    }
    const View& view() const {
        if (!m_view.ptr()) throw std::logic_error("Tensor::view() called on uninitialized tensor");
        return m_view;
    }
    View& view() {
        if (!m_view.ptr()) throw std::logic_error("Tensor::view() called on uninitialized tensor");
        return m_view;
    }
    T* ptr() const { return m_view.ptr(); }
    i64 size() const { return m_view.size(); }
    // flat element access (host tensors)
    T& operator[](i64 i) const { return m_view.ptr()[i]; }
    // new tensor of the same shape and location holding a copy of this one
    Tensor clone() const {
        Tensor t(m_view.m_memory);
        t.init(view().shape());
        m_view.to(t.view());
        return t;
    }
};

template <i32 N>
using MultiDim = std::vector<std::array<i64, N>>;

template <typename T, int Dim, Location L>
using MultiTensorView = std::vector<TensorView<T, Dim, L>>;

template <typename T, int Dim, Location L>
class MultiTensor {
   public:
    explicit MultiTensor(Memory* memory) : m_memory(memory) {}
    void init(const MultiDim<Dim>& shapes) {
        m_tensors.clear();
        m_tensors.reserve(shapes.size());
        m_views.resize(shapes.size());
        for (size_t i = 0; i < shapes.size(); i++) {
            m_tensors.emplace_back(m_memory);
            m_tensors[i].init(shapes[i]);
            m_views[i] = m_tensors[i].view();
        }
    }
    MultiTensorView<T, Dim, L>& views() { return m_views; }

   private:
    Memory* m_memory;
    std::vector<Tensor<T, Dim, L>> m_tensors;
    MultiTensorView<T, Dim, L> m_views;
};

#define VLM_ALIASES(T, S)                                                        \
    using TensorView1##S##D = TensorView<T, 1, Location::Device>;                \
    using TensorView2##S##D = TensorView<T, 2, Location::Device>;                \
    using TensorView3##S##D = TensorView<T, 3, Location::Device>;                \
    using TensorView1##S##H = TensorView<T, 1, Location::Host>;                  \
    using TensorView2##S##H = TensorView<T, 2, Location::Host>;                  \
    using TensorView3##S##H = TensorView<T, 3, Location::Host>;                  \
    using Tensor1##S##D = Tensor<T, 1, Location::Device>;                        \
    using Tensor2##S##D = Tensor<T, 2, Location::Device>;                        \
    using Tensor3##S##D = Tensor<T, 3, Location::Device>;                        \
    using Tensor3##S##H = Tensor<T, 3, Location::Host>;                          \
    using MultiTensor2##S##D = MultiTensor<T, 2, Location::Device>;              \
    using MultiTensor3##S##D = MultiTensor<T, 3, Location::Device>;              \
    using MultiTensor3##S##H = MultiTensor<T, 3, Location::Host>;                \
    using MultiTensorView2##S##D = MultiTensorView<T, 2, Location::Device>;      \
    using MultiTensorView3##S##D = MultiTensorView<T, 3, Location::Device>;      \
    using MultiTensorView3##S##H = MultiTensorView<T, 3, Location::Host>;

VLM_ALIASES(f32, f)
VLM_ALIASES(f64, d)
#undef VLM_ALIASES

}  // namespace vlm
