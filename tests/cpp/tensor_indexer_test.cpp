// Checks of bbfft::tensor_indexer (restating the cases of the reference's test/tensor.cpp).
#include "bbfft/tensor_indexer.hpp"

#include <array>
#include <cstdio>
#include <cstdlib>

using namespace bbfft;

static int failures = 0;
#define CHECK(x)                                                                                   \
    do {                                                                                           \
        if (!(x)) {                                                                                \
            std::printf("FAILED line %d: %s\n", __LINE__, #x);                                     \
            ++failures;                                                                            \
        }                                                                                          \
    } while (0)

template <std::size_t D> using A = std::array<std::size_t, D>;

int main() {
    { // col major packed
        auto x = tensor_indexer<std::size_t, 3u, layout::col_major>({3, 4, 5});
        CHECK(x(0, 0, 0) == 0);
        CHECK(x(1, 0, 0) == 1);
        CHECK(x(0, 1, 0) == 3);
        CHECK(x(0, 0, 1) == 12);
        CHECK(x(2, 2, 2) == 2 + 2 * 3 + 2 * 12);
        CHECK(x(2, 3, 4) == x.size() - 1);
        CHECK(x.size() == 60);
    }
    { // col major strided
        auto x = tensor_indexer<std::size_t, 3u, layout::col_major>({3, 4, 5}, {1, 5, 30});
        CHECK(x(0, 1, 0) == 5);
        CHECK(x(0, 0, 1) == 30);
        CHECK(x(2, 3, 4) == 137);
        CHECK(x.size() == 150);
    }
    { // may fuse / fused (col major)
        auto x1 = tensor_indexer<std::size_t, 5u, layout::col_major>({2, 3, 4, 5, 6});
        auto x2 = tensor_indexer<std::size_t, 5u, layout::col_major>({2, 3, 4, 5, 6},
                                                                     {1, 3, 9, 36, 216});
        CHECK(x1.may_fuse());
        CHECK((x1.may_fuse<1, 3>()));
        CHECK(x1.may_fuse<4>());
        CHECK(!x2.may_fuse());
        CHECK((x2.may_fuse<1, 3>()));
        CHECK((x2.may_fuse<2, 3>()));
        CHECK(!(x2.may_fuse<1, 4>()));
        auto f1 = x1.fused();
        auto f2 = x1.fused<1, 3>();
        CHECK(f1.dim() == 1 && f1.shape(0) == 720 && f1.stride(0) == 1);
        CHECK(f2.dim() == 3);
        CHECK(f2.shape() == (A<3>{2, 60, 6}));
        CHECK(f2.stride() == (A<3>{1, 2, 120}));
        CHECK(f2(1, 35, 3) == x1(1, 2, 3, 2, 3));
    }
    { // reshape mode (col major)
        auto x1 = tensor_indexer<std::size_t, 4u, layout::col_major>({10, 8, 5, 10},
                                                                     {1, 12, 120, 500});
        CHECK(x1.may_reshape_mode(1, A<3>{2, 2, 2}));
        CHECK(!x1.may_reshape_mode(1, A<3>{2, 3, 2}));
        auto r1 = x1.reshaped_mode(0, A<2>{2, 5});
        CHECK(r1.dim() == 5u && r1.shape() == (A<5>{2, 5, 8, 5, 10}));
        CHECK(r1.stride() == (A<5>{1, 2, 12, 120, 500}));
        auto r2 = x1.reshaped_mode(1, A<3>{2, 2, 2});
        CHECK(r2.shape() == (A<6>{10, 2, 2, 2, 5, 10}));
        CHECK(r2.stride() == (A<6>{1, 12, 24, 48, 120, 500}));
        auto r3 = x1.reshaped_mode(3, A<2>{5, 2});
        CHECK(r3.shape() == (A<5>{10, 8, 5, 5, 2}));
        CHECK(r3.stride() == (A<5>{1, 12, 120, 500, 2500}));
    }
    { // row major packed / strided
        auto x = tensor_indexer<std::size_t, 3u, layout::row_major>({3, 4, 5});
        CHECK(x(1, 0, 0) == 20);
        CHECK(x(0, 1, 0) == 5);
        CHECK(x(0, 0, 1) == 1);
        CHECK(x(2, 3, 4) == x.size() - 1);
        auto y = tensor_indexer<std::size_t, 3u, layout::row_major>({3, 4, 5}, {30, 6, 1});
        CHECK(y(2, 3, 4) == 82);
        CHECK(y.size() == 90);
    }
    { // row major fuse / reshape
        auto x1 = tensor_indexer<std::size_t, 5u, layout::row_major>({6, 5, 4, 3, 2});
        auto x2 = tensor_indexer<std::size_t, 5u, layout::row_major>({6, 5, 4, 3, 2},
                                                                     {216, 36, 9, 3, 1});
        CHECK(!x2.may_fuse());
        CHECK((x2.may_fuse<1, 3>()));
        CHECK(!(x2.may_fuse<1, 4>()));
        auto f2 = x1.fused<1, 3>();
        CHECK(f2.shape() == (A<3>{6, 60, 2}));
        CHECK(f2.stride() == (A<3>{120, 2, 1}));
        CHECK(f2(1, 35, 3) == x1(1, 2, 3, 2, 3));
        auto x3 = tensor_indexer<std::size_t, 4u, layout::row_major>({10, 8, 5, 10});
        auto r1 = x3.reshaped_mode(0, A<2>{2, 5});
        CHECK(r1.shape() == (A<5>{2, 5, 8, 5, 10}));
        CHECK(r1.stride() == (A<5>{2000, 400, 50, 10, 1}));
        auto r2 = x3.reshaped_mode(1, A<3>{2, 2, 2});
        CHECK(r2.stride() == (A<6>{400, 200, 100, 50, 10, 1}));
        auto r3 = x3.reshaped_mode(3, A<2>{5, 2});
        CHECK(r3.stride() == (A<5>{400, 50, 10, 2, 1}));
    }
    { // fit_array
        auto a = fit_array<5>(A<3>{1, 2, 3});
        CHECK(a == (A<5>{1, 2, 3, 0, 0}));
        auto b = fit_array<2>(A<3>{1, 2, 3});
        CHECK(b == (A<2>{1, 2}));
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
