// mat-gpu -- command-line front end with the reference's flags
// (mat-gpu/main.cu:220-533): -h -i -n K,M,N -b nbits -m method -p,
// plus -s seed (the reference seeds with the time, main.cu:342) and
// -j (one JSON line with the timing on stdout after the product).
//
// Methods (main.cu:358-502): 1 and 2 naive (1 ran the arithmetic on the host
// in the reference; here both use the device kernel), 3 Strassen, 35
// Winograd, 4 naive recursion (squarify), 37 / 357 / 47 the same on the
// interleaved layout, and new: 8 the INT8-MFMA CRT route.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>
#include "mgx.h"
#include "mgx_sched.h"

static int usage(const char *argv0)
{
    fprintf(stderr,
            "Usage: %s [-hip] [-n size] [-b nbits] [-m method] [-s seed] [-j] [-M model]\n"
            "  -h         Show help\n"
            "  -i         Read matrices from input instead of generating randomly\n"
            "  -n K,M,N   Set the size for random matrix generation (KxM * MxN = KxN)\n"
            "  -b nbits   Number of bits for generated numbers\n"
            "  -m method  1/2 naive, 3 Strassen, 35 Winograd, 4 naive recursive,\n"
            "             37/357/47 interleaved Strassen/Winograd/naive, 8 CRT-MFMA,\n"
            "             0 the route the cost model picks\n"
            "  -p         Print result after multiplication\n"
            "  -s seed    Seed for generated numbers (default: the time)\n"
            "  -j         Print a JSON line with the timing\n"
            "  -M model   Cost model for -m 0 (default: costmodel_params.txt beside bin/)\n",
            argv0);
    return 0;
}

static bool read_token(std::string &tok)
{
    tok.clear();
    int c;
    while ((c = getchar()) != EOF && isspace(c)) {}
    if (c == EOF) return false;
    do tok.push_back((char)c);
    while ((c = getchar()) != EOF && !isspace(c));
    return true;
}

int main(int argc, char **argv)
{
    bool read_matrices = false, print_result = false, json = false;
    long long K = -1, M = -1, N = -1, nbits = -1;
    int method = 1;
    unsigned long long seed = (unsigned long long)time(nullptr);
    std::string model_path;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (a[0] != '-' || !a[1]) { fprintf(stderr, "Unknown flag, use -h to get info\n"); return 1; }
        for (int j = 1; a[j]; j++) {
            char f = a[j];
            auto value = [&]() -> const char * {
                if (a[j + 1]) { const char *v = a + j + 1; j = (int)strlen(a) - 1; return v; }
                if (i + 1 < argc) return argv[++i];
                return nullptr;
            };
            if (f == 'h') return usage(argv[0]);
            else if (f == 'i') read_matrices = true;
            else if (f == 'p') print_result = true;
            else if (f == 'j') json = true;
            else if (f == 'n') {
                const char *v = value();
                if (!v || sscanf(v, "%lld,%lld,%lld", &K, &M, &N) != 3) {
                    fprintf(stderr, "Invalid number of commas in argument for -n\n");
                    return 1;
                }
                break;
            } else if (f == 'b') { const char *v = value(); if (!v) return 1; nbits = atoll(v); break; }
            else if (f == 'm') { const char *v = value(); if (!v) return 1; method = atoi(v); break; }
            else if (f == 's') { const char *v = value(); if (!v) return 1; seed = strtoull(v, nullptr, 10); break; }
            else if (f == 'M') { const char *v = value(); if (!v) return 1; model_path = v; break; }
            else { fprintf(stderr, "Unknown flag, use -h to get info\n"); return 1; }
        }
    }
    std::vector<std::vector<std::string>> textA, textB;
    if (read_matrices) {
        if (K != -1) { fprintf(stderr, "Cannot specify matrix size if reading matrices from input\n"); return 1; }
        if (nbits != -1) { fprintf(stderr, "Cannot specify nbits if reading matrices from input\n"); return 1; }
        if (scanf("%lld %lld %lld", &K, &M, &N) != 3) { fprintf(stderr, "Invalid matrix header\n"); return 1; }
    } else if (nbits == -1) {
        fprintf(stderr, "No bit count specified!\n");
        return 1;
    }
    if (K < 0 || M < 0 || N < 0) { fprintf(stderr, "No matrix size specified!\n"); return 1; }

    // capacity: the product of two entries summed M times, plus room for
    // the recursion's operand growth
    long long inbits = nbits;
    std::vector<std::string> tokA, tokB;
    if (read_matrices) {
        std::string t;
        inbits = 1;
        for (long long i = 0; i < K * M + M * N; i++) {
            if (!read_token(t)) { fprintf(stderr, "Unexpected end of input\n"); return 1; }
            size_t digits = t.size() - (t[0] == '-');
            inbits = std::max<long long>(inbits, (long long)digits * 4);
            (i < K * M ? tokA : tokB).push_back(t);
        }
    }
    long long maxdim = std::max(std::max(K, M), std::max(N, 1LL));
    long long cap = (2 * inbits + (long long)std::ceil(std::log2((double)maxdim)) + 1 + 16 + 63) / 64 + 1;
    if (mgx_init(0) != MGX_OK) { fprintf(stderr, "mgx_init: %s\n", mgx_last_error()); return 1; }
    const long long W = cap + 1;
    std::vector<mgx_limb> A((size_t)std::max(K * M, 1LL) * W), B((size_t)std::max(M * N, 1LL) * W),
        C((size_t)std::max(K * N, 1LL) * W);
    if (read_matrices) {
        for (long long i = 0; i < K * M; i++)
            if (mgx_entry_from_hex(&A[i * W], cap, tokA[i].c_str()) != MGX_OK) {
                fprintf(stderr, "%s\n", mgx_last_error());
                return 1;
            }
        for (long long i = 0; i < M * N; i++)
            if (mgx_entry_from_hex(&B[i * W], cap, tokB[i].c_str()) != MGX_OK) {
                fprintf(stderr, "%s\n", mgx_last_error());
                return 1;
            }
    } else {
        fprintf(stderr, "Seed %llu\n", seed);
        if (mgx_random_matrix(A.data(), M, K, M, cap, nbits, seed) != MGX_OK ||
            mgx_random_matrix(B.data(), N, M, N, cap, nbits, seed + 1) != MGX_OK) {
            fprintf(stderr, "%s\n", mgx_last_error());
            return 1;
        }
    }
    mgx_costmodel *model = nullptr;
    if (method == 0) {   // learned dispatch: install the model, else the autotuner times the routes
        if (model_path.empty()) {
            std::string exe = argv[0];
            size_t slash = exe.rfind('/');
            model_path = (slash == std::string::npos ? std::string(".") : exe.substr(0, slash)) +
                         "/../costmodel_params.txt";
        }
        if (mgx_costmodel_load(&model, model_path.c_str()) == MGX_OK) mgx_dispatch_set_model(model);
        else fprintf(stderr, "No cost model (%s); timing the routes instead\n", mgx_last_error());
    }
    fprintf(stderr, "Using method %d\n", method);
    double secs = 0;
    int st = mgx_matmul_host(method, K, M, N, C.data(), A.data(), B.data(), cap, &secs);
    if (model) {
        mgx_dispatch_set_model(nullptr);
        mgx_costmodel_destroy(model);
    }
    if (st != MGX_OK) { fprintf(stderr, "Multiplication failed: %s\n", mgx_last_error()); return 1; }
    fprintf(stderr, "Time taken: %g seconds\n", secs);
    if (print_result) {
        std::vector<char> buf((size_t)cap * 16 + 4);
        for (long long y = 0; y < K; y++) {
            for (long long x = 0; x < N; x++) {
                if (x) putchar(' ');
                mgx_entry_to_hex(buf.data(), (int64_t)buf.size(), &C[(y * N + x) * W], cap);
                fputs(buf.data(), stdout);
            }
            putchar('\n');
        }
        putchar('\n');
    }
    if (json)
        printf("{\"method\": %d, \"K\": %lld, \"M\": %lld, \"N\": %lld, \"seconds\": %.9g}\n", method, K, M, N, secs);
    return 0;
}
