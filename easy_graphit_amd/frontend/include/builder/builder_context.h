// builder_context::extract_function_ast: stage a C++ function whose dyn_var
// parameters become the generated function's parameters (arg0, arg1, ...)
// and whose other parameters are bound to the given static arguments, in
// order.
#pragma once

#include <tuple>

#include "builder/dyn_var.h"

namespace builder {

template <class T>
struct is_dyn : std::false_type {};
template <class T>
struct is_dyn<dyn_var<T>> : std::true_type {};
template <class T>
struct dyn_value;
template <class T>
struct dyn_value<dyn_var<T>> {
    using type = T;
};

template <class... P>
struct param_kinds {
    static constexpr bool dyn[sizeof...(P) + 1] = {is_dyn<std::decay_t<P>>::value..., false};
    static constexpr int statics_before(int i) {
        int c = 0;
        for (int j = 0; j < i; ++j) c += dyn[j] ? 0 : 1;
        return c;
    }
    static constexpr int dyns_before(int i) {
        int c = 0;
        for (int j = 0; j < i; ++j) c += dyn[j] ? 1 : 0;
        return c;
    }
};

class builder_context {
  public:
    template <class R, class... P, class... S>
    block::block::Ptr extract_function_ast(R (*fn)(P...), const std::string &name, S &&...statics) {
        static_assert(sizeof...(S) == (0 + ... + (is_dyn<std::decay_t<P>>::value ? 0 : 1)),
                      "one static argument per non-dyn_var parameter");
        auto f = std::make_shared<block::function>();
        f->name = name;
        auto st = std::forward_as_tuple(std::forward<S>(statics)...);
        context ctx;
        push_context(&ctx);
        try {
            f->body = extract_body([&]() { call_with(fn, st, std::index_sequence_for<P...>{}); });
        } catch (...) {
            pop_context();
            throw;
        }
        pop_context();
        add_params<P...>(*f, std::index_sequence_for<P...>{});
        return f;
    }

  private:
    template <class... P, size_t... I>
    static void add_params(block::function &f, std::index_sequence<I...>) {
        (add_param<P, I, P...>(f), ...);
    }
    template <class PI, size_t I, class... P>
    static void add_param(block::function &f) {
        if constexpr (is_dyn<std::decay_t<PI>>::value) {
            using T = typename dyn_value<std::decay_t<PI>>::type;
            f.params.emplace_back(type_name<T>::get(), "arg" + std::to_string(param_kinds<P...>::dyns_before(I)));
        }
    }
    template <size_t I, class PI, class... P, class Tup>
    static decltype(auto) arg(Tup &st) {
        if constexpr (is_dyn<std::decay_t<PI>>::value) {
            using T = typename dyn_value<std::decay_t<PI>>::type;
            block::ptr v = block::make(block::kind::var, "arg" + std::to_string(param_kinds<P...>::dyns_before(I)));
            v->type = type_name<T>::get();
            return dyn_var<T>(alias, v);
        } else {
            return std::get<param_kinds<P...>::statics_before(I)>(st);
        }
    }
    template <class R, class... P, class Tup, size_t... I>
    static void call_with(R (*fn)(P...), Tup &st, std::index_sequence<I...>) {
        fn(arg<I, P, P...>(st)...);
    }
};

}  // namespace builder
