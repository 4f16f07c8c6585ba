// ut_anderson -- Anderson acceleration (vlm_solvers) over BackendHIP on a linear contraction
// x = M x + c with M_ij = 0.9 sin(1 + i + 3 j) / n, c_i = cos(i), n = 64.  Prints the iteration
// count and the fixed point: "ITER <k>" then "X <i> <x_i>" lines (FP64), then the same for FP32
// with prefix F32.  Usage: ut_anderson [m] [tol64] [tol32]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vlm_solvers.hpp"

using namespace vlm;

template <typename T>
static int run(Backend& b, int m, T tol, const char* tag) {
    const i64 n = 64;
    Memory* mem = b.memory.get();
    Tensor<T, 2, Location::Host> Mh{mem};
    Tensor<T, 1, Location::Host> ch{mem}, xh{mem};
    Tensor<T, 2, Location::Device> Md{mem};
    Tensor<T, 1, Location::Device> cd{mem}, xd{mem};
    Mh.init({n, n});
    ch.init({n});
    xh.init({n});
    Md.init({n, n});
    cd.init({n});
    xd.init({n});
    for (i64 j = 0; j < n; j++)
        for (i64 i = 0; i < n; i++) Mh.view()(i, j) = (T)(0.9 * std::sin(1.0 + i + 3.0 * j) / n);
    for (i64 i = 0; i < n; i++) ch.view()(i) = (T)std::cos((double)i);
    Mh.view().to(Md.view());
    ch.view().to(cd.view());
    xd.view().fill(0);
    auto f = [&](const TensorView<T, 1, Location::Device>& x, const TensorView<T, 1, Location::Device>& y) {
        cd.view().to(y);
        b.blas->gemv((T)1, Md.view(), x, (T)1, y);
    };
    const i32 k = anderson_acceleration(&b, xd.view(), f, 100, tol, m);
    xd.view().to(xh.view());
    std::printf("%sITER %d\n", tag, k);
    for (i64 i = 0; i < n; i++) std::printf("%sX %lld %.17g\n", tag, (long long)i, (double)xh.view()(i));
    return k < 100 ? 0 : 1;
}

int main(int argc, char** argv) {
    const int m = argc > 1 ? std::atoi(argv[1]) : 3;
    const double tol64 = argc > 2 ? std::atof(argv[2]) : 1e-12;
    const float tol32 = argc > 3 ? (float)std::atof(argv[3]) : 1e-5f;
    try {
        auto b = create_backend("hip");
        int rc = run<double>(*b, m, tol64, "");
        rc |= run<float>(*b, m, tol32, "F32");
        return rc;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 3;
    }
}
