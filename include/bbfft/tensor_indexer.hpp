// Address computation for strided multi-dimensional tensors (host utility used by tests/users).
// Mirrors the reference interface include/bbfft/tensor_indexer.hpp:24-354: same class template,
// member names and semantics (including that may_fuse/fused take mode numbers in memory order,
// fastest mode first, for both layouts).
#ifndef BBFFT_TENSOR_INDEXER_HPP
#define BBFFT_TENSOR_INDEXER_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <utility>

namespace bbfft {

/// row_major: last index varies fastest; col_major: first index varies fastest.
enum class layout { row_major, col_major };

/// Copy the first min(Din, Dout) entries of `in` and pad the rest with fill_value.
template <std::size_t Dout, typename IdxT, std::size_t Din>
constexpr auto fit_array(std::array<IdxT, Din> const &in, IdxT fill_value = IdxT(0)) {
    std::array<IdxT, Dout> out{};
    for (std::size_t d = 0; d < Dout; ++d) {
        out[d] = d < Din ? in[d] : fill_value;
    }
    return out;
}

template <typename IdxT, unsigned int D, layout L = layout::row_major> class tensor_indexer {
  public:
    using multi_idx_t = std::array<IdxT, D>;

    tensor_indexer() : mshape_{}, mstride_{} {}

    /// Packed tensor.
    tensor_indexer(multi_idx_t shape) : mshape_(to_mem(shape)) {
        IdxT s = IdxT(1);
        for (unsigned i = 0; i < D; ++i) {
            mstride_[i] = s;
            s *= mshape_[i];
        }
    }
    /// Tensor with explicit strides (given in the same index order as shape).
    tensor_indexer(multi_idx_t shape, multi_idx_t stride)
        : mshape_(to_mem(shape)), mstride_(to_mem(stride)) {}

    template <typename... Indices, typename = std::enable_if_t<sizeof...(Indices) == D, int>>
    IdxT operator()(Indices &&...is) const {
        multi_idx_t idx = {static_cast<IdxT>(is)...};
        return (*this)(idx);
    }
    IdxT operator()(multi_idx_t const &idx) const {
        IdxT r = IdxT(0);
        for (unsigned i = 0; i < D; ++i) {
            r += idx[i] * stride(i);
        }
        return r;
    }

    auto shape() const { return to_mem(mshape_); } // to_mem is its own inverse
    auto shape(unsigned int d) const { return mshape_[mem_pos(d)]; }
    auto stride() const { return to_mem(mstride_); }
    auto stride(unsigned int d) const { return mstride_[mem_pos(d)]; }
    /// Number of elements spanned (slowest stride times slowest extent).
    IdxT size() const { return mstride_[D - 1] * mshape_[D - 1]; }
    constexpr auto dim() const { return D; }

    /// Modes Dfrom..Dto (memory order) are packed and may be treated as one mode.
    template <unsigned int Dfrom = 0, unsigned int Dto = D - 1> bool may_fuse() const {
        static_assert(Dfrom <= Dto && Dto <= D - 1);
        for (unsigned d = Dfrom; d < Dto; ++d) {
            if (mstride_[d] * mshape_[d] != mstride_[d + 1]) {
                return false;
            }
        }
        return true;
    }

    /// Indexer with modes Dfrom..Dto (memory order) merged into one.
    template <unsigned int Dfrom = 0, unsigned int Dto = D - 1> auto fused() const {
        static_assert(Dfrom <= Dto && Dto <= D - 1);
        constexpr unsigned E = D - (Dto - Dfrom);
        std::array<IdxT, E> s{}, st{};
        unsigned o = 0;
        for (unsigned d = 0; d < D; ++d) {
            if (d > Dfrom && d <= Dto) {
                s[o - 1] *= mshape_[d];
                continue;
            }
            s[o] = mshape_[d];
            st[o] = mstride_[d];
            ++o;
        }
        if constexpr (L == layout::row_major) {
            std::reverse(s.begin(), s.end());
            std::reverse(st.begin(), st.end());
        }
        return tensor_indexer<IdxT, E, L>(s, st);
    }

    /// Mode `mode` (index order of this indexer) has as many entries as mode_shape spans.
    template <std::size_t E>
    bool may_reshape_mode(int mode, std::array<IdxT, E> const &mode_shape) const {
        IdxT n = IdxT(1);
        for (auto v : mode_shape) {
            n *= v;
        }
        return n == shape(mode);
    }

    /// View mode `mode` as an E-mode sub-tensor of shape mode_shape (index order of this indexer).
    template <std::size_t E> auto reshaped_mode(int mode, std::array<IdxT, E> mode_shape) const {
        static_assert(E > 0u);
        constexpr unsigned F = D + E - 1;
        unsigned const mp = mem_pos(mode);
        if constexpr (L == layout::row_major) {
            std::reverse(mode_shape.begin(), mode_shape.end());
        }
        std::array<IdxT, F> s{}, st{};
        unsigned o = 0;
        for (unsigned d = 0; d < D; ++d) {
            if (d == mp) {
                IdxT cur = mstride_[d];
                for (std::size_t e = 0; e < E; ++e) {
                    s[o] = mode_shape[e];
                    st[o] = cur;
                    cur *= mode_shape[e];
                    ++o;
                }
            } else {
                s[o] = mshape_[d];
                st[o] = mstride_[d];
                ++o;
            }
        }
        if constexpr (L == layout::row_major) {
            std::reverse(s.begin(), s.end());
            std::reverse(st.begin(), st.end());
        }
        return tensor_indexer<IdxT, F, L>(s, st);
    }

  private:
    static constexpr unsigned mem_pos(unsigned d) {
        return L == layout::row_major ? D - 1 - d : d;
    }
    static multi_idx_t to_mem(multi_idx_t a) {
        if constexpr (L == layout::row_major) {
            std::reverse(a.begin(), a.end());
        }
        return a;
    }

    multi_idx_t mshape_;  // memory order: fastest mode first
    multi_idx_t mstride_; // memory order
};

template <typename IdxT, layout L = layout::row_major>
using vector_indexer = tensor_indexer<IdxT, 1u, L>;
template <typename IdxT, layout L = layout::row_major>
using matrix_indexer = tensor_indexer<IdxT, 2u, L>;

} // namespace bbfft

#endif // BBFFT_TENSOR_INDEXER_HPP
