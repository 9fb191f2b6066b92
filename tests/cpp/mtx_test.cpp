// Matrix-Market readers of the C++ drop-in (BlockSp::tests::mtxMatrix_to_BlockSp and
// mtxArray_to_solArray, blocksp_tests.hpp:595-699) on the host: reads one file and prints
// what it read, for tests/test_mtx.py to compare with its own statement of the format.
//
//   mtx_test mat  <path> <start_line> <symmetric> <bs>   bs in {1, 2, 4, 8}, double
//   mtx_test cmat <path> <start_line> <symmetric>        std::complex<double>, bs 2
//   mtx_test vec  <path> <start_line>                    double, B 2, D 2
//   mtx_test cvec <path> <start_line>                    std::complex<double>, B 2, D 1
//
// Matrix output: "m n", then per block row "ncols c0 c1 ...", then every block's entries
// row-major in the row's order, one per line ("re" or "re im", %.17g).  Vector output: the
// flat values in solArray order.  A BlockSp::Error prints "ERROR <message>" and exits 3.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "blocksp_include.hpp"

namespace {

void put(double v) { std::printf("%.17g\n", v); }
void put(const std::complex<double>& v) { std::printf("%.17g %.17g\n", v.real(), v.imag()); }

template <typename T, int B>
void dump(const BlockSp::BlockSp<T, B>& A) {
    std::printf("%d %d\n", A.m(), A.n());
    for (int im = 0; im < A.m(); ++im) {
        std::printf("%zu", A.colSize(im));
        for (int c : A.get_colIndex(im)) std::printf(" %d", c);
        std::printf("\n");
    }
    for (int im = 0; im < A.m(); ++im)
        for (const auto& blk : A.get_data(im))
            for (int r = 0; r < B; ++r)
                for (int c = 0; c < B; ++c) put(blk(r, c));
}

template <int B>
void read_mat(const char* path, unsigned start, bool sym) {
    dump(BlockSp::tests::mtxMatrix_to_BlockSp<double, B>(path, start, sym));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: mtx_test mat|cmat|vec|cvec <path> <start_line> [symmetric] [bs]\n");
        return 2;
    }
    const std::string mode = argv[1];
    const char* path = argv[2];
    const unsigned start = (unsigned)std::atoi(argv[3]);
    const bool sym = argc > 4 && std::atoi(argv[4]) != 0;
    try {
        if (mode == "mat") {
            switch (argc > 5 ? std::atoi(argv[5]) : 1) {
                case 1: read_mat<1>(path, start, sym); break;
                case 2: read_mat<2>(path, start, sym); break;
                case 4: read_mat<4>(path, start, sym); break;
                case 8: read_mat<8>(path, start, sym); break;
                default: std::fprintf(stderr, "bs must be 1, 2, 4 or 8\n"); return 2;
            }
        } else if (mode == "cmat") {
            dump(BlockSp::tests::mtxMatrix_to_BlockSp<std::complex<double>, 2>(path, start, sym));
        } else if (mode == "vec") {
            const auto x = BlockSp::tests::mtxArray_to_solArray<double, 2, 2>(path, start);
            for (size_t i = 0; i < x.size_flat(); ++i) put(x.data()[i]);
        } else if (mode == "cvec") {
            const auto x = BlockSp::tests::mtxArray_to_solArray<std::complex<double>, 2, 1>(path, start);
            for (size_t i = 0; i < x.size_flat(); ++i) put(x.data()[i]);
        } else {
            std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
            return 2;
        }
    } catch (const BlockSp::Error& e) {
        std::printf("ERROR %s\n", e.what());
        return 3;
    }
    return 0;
}
