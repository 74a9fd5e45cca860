// Known-answer generator for the MT19937-64 restatement: prints the first n
// outputs of std::mt19937_64 seeded with argv[1] (the C++ standard engine has
// the same recurrence and seeding as init_genrand64/genrand64_int64).
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
    unsigned long long seed = std::strtoull(argv[1], nullptr, 10);
    long n = std::atol(argv[2]);
    std::mt19937_64 e(seed);
    for (long i = 0; i < n; i++) std::printf("%llu\n", (unsigned long long)e());
    return 0;
}
