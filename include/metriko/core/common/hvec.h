// metriko (MI355X build) -- host vector type of the large CSR and operator arrays.
#ifndef METRIKO_HVEC_H
#define METRIKO_HVEC_H
#include <complex>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace metriko {
    template<typename T> struct IsComplex : std::false_type {};
    template<typename T> struct IsComplex<std::complex<T>> : std::true_type {};

    // Allocator whose resize() leaves new scalars uninitialised (default-initialised): the
    // large index and value arrays are then first written, and so first touched, by the
    // parallel loops that fill them, instead of being zeroed by one thread beforehand.
    // std::complex counts as a scalar here: its default constructor would write zeros (it is
    // trivially copyable and destructible, so the storage needs no constructor call).
    template<typename T>
    struct DefaultInitAllocator : std::allocator<T> {
        template<typename U> struct rebind { using other = DefaultInitAllocator<U>; };
        DefaultInitAllocator() = default;
        template<typename U> DefaultInitAllocator(const DefaultInitAllocator<U>&) noexcept {}
        template<typename U> void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
            if constexpr (!IsComplex<U>::value) ::new (static_cast<void*>(p)) U;
        }
        template<typename U, typename... A> void construct(U* p, A&&... a) {
            ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
        }
    };
    // Vector type of the CSR arrays: resize() does not zero (every caller fills what it sizes).
    template<typename T>
    using HVec = std::vector<T, DefaultInitAllocator<T>>;

}

#endif
