// ut_tensor -- slicing, copy, fill and reshape of the tensor data model (the cases of the
// reference's tests/ut_tensor.cpp, with the expected values worked out from column-major
// addressing).  `ut_tensor` runs them on host tensors only (no GPU needed); `ut_tensor --device`
// repeats the copy cases through device tensors of the HIP backend.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "vlm_backend.hpp"
#include "vlm_memory.hpp"

using namespace vlm;

static int g_fail = 0;
#define CHECK_EQ(a, b)                                                                                   \
    do {                                                                                                 \
        const auto va_ = (a);                                                                            \
        const auto vb_ = (b);                                                                            \
        if (!(va_ == vb_)) {                                                                             \
            std::printf("%s:%d: %s == %s failed (%g vs %g)\n", __FILE__, __LINE__, #a, #b, (double)va_, \
                        (double)vb_);                                                                    \
            g_fail++;                                                                                    \
        }                                                                                                \
    } while (0)

template <size_t N>
static void check_strides(const std::array<i64, N>& got, const std::array<i64, N>& want, int line) {
    for (size_t i = 0; i < N; i++)
        if (got[i] != want[i]) {
            std::printf("ut_tensor.cpp:%d: stride[%zu] = %lld, expected %lld\n", line, i, (long long)got[i],
                        (long long)want[i]);
            g_fail++;
        }
}

class HostMemory final : public Memory {
   public:
    HostMemory() : Memory(true) {}
    void* alloc(Location, i64 bytes) const override { return std::malloc((size_t)bytes); }
    void free(Location, void* p) const override { std::free(p); }
    void copy(Location, void* dst, i64 ds, Location, const void* src, i64 ss, i64 es, i64 n) const override {
        for (i64 i = 0; i < n; i++) std::memcpy((char*)dst + i * ds * es, (const char*)src + i * ss * es, (size_t)es);
    }
    void fill(Location, float* p, i64 s, float v, i64 n) const override {
        for (i64 i = 0; i < n; i++) p[i * s] = v;
    }
    void fill(Location, double* p, i64 s, double v, i64 n) const override {
        for (i64 i = 0; i < n; i++) p[i * s] = v;
    }
};

// work tensor W (3x3x4 holding 0..35) at location L; results copied into the host tensor out
template <Location L>
static void run_cases(Memory* memory) {
    Tensor<f32, 3, Location::Host> base{memory};
    base.init({3, 3, 4});
    for (u32 i = 0; i < base.size(); i++) base[i] = (f32)i;
    Tensor<f32, 3, L> work{memory};
    work.init({3, 3, 4});
    base.view().to(work.view());
    Tensor<f32, 3, Location::Host> out{memory};
    out.init({3, 3, 4});
    const auto& o = out.view();

    {  // strided host view: element (i, 1, k) = i + 3 + 9k
        auto bv = base.view().slice(All, 1, Range{0, 3});
        CHECK_EQ(bv(0, 0), 3.0f);
        CHECK_EQ(bv(1, 0), 4.0f);
        CHECK_EQ(bv(2, 1), 14.0f);
        CHECK_EQ(bv(2, 2), 23.0f);
    }
    {  // (i, 0..1, 0) -> (i, 1..2, 2)
        auto t = work.clone();
        const auto& tv = t.view();
        tv.slice(All, Range{0, 2}, 0).to(tv.slice(All, Range{1, 3}, 2));
        tv.to(o);
        CHECK_EQ(o(0, 2, 2), 3.0f);
        CHECK_EQ(o(1, 1, 2), 1.0f);
        CHECK_EQ(o(2, 2, 2), 5.0f);
        CHECK_EQ(o(0, 0, 2), 18.0f);  // untouched
    }
    {  // row 0 -> row 1
        auto t = work.clone();
        const auto& tv = t.view();
        tv.slice(0, All, All).to(tv.slice(1, All, All));
        tv.to(o);
        CHECK_EQ(o(0, 0, 0), o(1, 0, 0));
        CHECK_EQ(o(0, 1, 1), o(1, 1, 1));
        CHECK_EQ(o(1, 2, 3), 33.0f);
    }
    {  // nested slices of orthogonal planes
        auto t = work.clone();
        const auto& tv = t.view();
        auto a = tv.slice(0, All, All).slice(Range{0, 2}, All);
        auto b = tv.slice(All, 2, All).slice(Range{1, 3}, All);
        a.to(b);
        tv.to(o);
        CHECK_EQ(o(0, 0, 1), o(1, 2, 1));
        CHECK_EQ(o(0, 0, 2), o(1, 2, 2));
        CHECK_EQ(o(2, 2, 3), 30.0f);  // W(0,1,3) = 3 + 27
    }
    {  // orthogonal slices: (0, j, k<2) -> (j, 1+k, 2)
        auto t = work.clone();
        const auto& tv = t.view();
        tv.slice(0, All, Range{0, 2}).to(tv.slice(All, Range{1, 3}, 2));
        tv.to(o);
        CHECK_EQ(o(0, 0, 0), o(0, 1, 2));
        CHECK_EQ(o(0, 1, 0), o(1, 1, 2));
        CHECK_EQ(o(0, 2, 1), o(2, 2, 2));
        CHECK_EQ(o(2, 2, 2), 15.0f);
    }
    {  // fill one plane
        auto t = work.clone();
        const auto& tv = t.view();
        tv.slice(0, All, All).fill(111.f);
        tv.to(o);
        CHECK_EQ(o(0, 0, 0), 111.f);
        CHECK_EQ(o(0, 1, 1), 111.f);
        CHECK_EQ(o(0, 2, 2), 111.f);
        CHECK_EQ(o(1, 2, 2), 25.0f);
    }
    {  // negative indices count from the end
        CHECK_EQ(base.view()(-1, -1, -1), 35.0f);
        CHECK_EQ(base.view().slice(All, -1, All)(0, 1), 15.0f);
    }
}

static void reshape_cases(Memory* memory) {
    Tensor<f32, 3, Location::Host> t{memory};
    t.init({3, 3, 4});
    const auto& tv = t.view();
    check_strides(tv.slice(Range{0, 2}, Range{1, 3}, All).reshape(1, 2, 2, 2, 2).stride(),
                  std::array<i64, 5>{1, 1, 3, 9, 18}, __LINE__);
    check_strides(tv.slice(All, Range{1, 3}, All).reshape(2, 3, 1, 1, 4).stride(), std::array<i64, 5>{1, 2, 9, 9, 9},
                  __LINE__);
    check_strides(tv.slice(0, All, All).reshape(3, 2, 2).stride(), std::array<i64, 3>{3, 9, 18}, __LINE__);
    check_strides(tv.reshape(2, 18).stride(), std::array<i64, 2>{1, 2}, __LINE__);
    check_strides(tv.reshape(36).stride(), std::array<i64, 1>{1}, __LINE__);
    Tensor<f32, 1, Location::Host> v{memory};
    v.init({10});
    check_strides(v.view().reshape(10, 1, 1).stride(), std::array<i64, 3>{1, 10, 10}, __LINE__);
    bool threw = false;
    try {
        (void)tv.slice(Range{0, 2}, All, All).reshape(24);  // rows 0..1 of each column: not one run
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK_EQ(threw, true);
}

int main(int argc, char** argv) {
    const bool device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
    HostMemory host;
    run_cases<Location::Host>(&host);
    reshape_cases(&host);
    if (device) {
        for (const auto& name : get_available_backends()) {
            auto backend = create_backend(name);
            auto memory = backend->create_memory_manager();
            run_cases<Location::Device>(memory.get());
            reshape_cases(memory.get());
            std::printf("backend %s: device cases run\n", name.c_str());
        }
    }
    if (g_fail) std::printf("ut_tensor: %d failures\n", g_fail);
    else std::printf("ut_tensor: all passed\n");
    return g_fail ? 1 : 0;
}
