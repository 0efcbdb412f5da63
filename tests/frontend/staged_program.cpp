// A staged program of this project's own (not one of the reference's apps)
// for the front end's CPU tests: loops with a condition that is a plain
// expression, a nested counted loop, an if/else inside a loop, a loop whose
// condition declares a variable each time it is tested, and a branch after
// the loops. The generator prints it through run_graphit_pipeline.
#include "builder/builder_context.h"
#include "builder/dyn_var.h"
#include "pipeline/graphit.h"

dyn_var<int> result("result");

static void program(dyn_var<int> n, dyn_var<int> m) {
    dyn_var<int> i = 0;
    while (i < n) {
        if (i % 3 == 0)
            m = m + i;
        else
            m = m - 1;
        for (dyn_var<int> j = 0; j < 2; j = j + 1) m = (m * 2) % 1000003;
        i = i + 1;
    }
    dyn_var<int> k = m % 7;
    while (dyn_var<int>(k > 0)) {
        m = m + k;
        k = k - 1;
    }
    if (m < 0)
        result = -m;
    else
        result = m;
}

int main() {
    auto ast = builder::builder_context().extract_function_ast(program, "program");
    pipeline::run_graphit_pipeline(ast, std::cout);
    return 0;
}
