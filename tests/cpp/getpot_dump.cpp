// Prints every variable tk::getpot reads from a file (argv[1]) as "name<TAB>value", then the
// lookups the loader makes under the prefix argv[2] for the names argv[3..] as int, double
// and string.  Test helper of tests/test_getpot.py (host code only).
#include <cstdio>

#include "../../include/tk/getpot.hxx"

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    tk::getpot g(argv[1], "#", "\n");
    for (auto const& kv : g.variables()) std::printf("var\t%s\t%s\n", kv.first.c_str(), kv.second.c_str());
    g.set_prefix(argv[2]);
    for (int a = 3; a < argc; ++a)
        std::printf("get\t%s\t%d\t%.17g\t%s\n", argv[a], g(argv[a], -1), g(argv[a], -1.),
                    g(argv[a], "").c_str());
    return 0;
}
