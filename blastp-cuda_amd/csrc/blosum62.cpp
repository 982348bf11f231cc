// blosum62.cpp -- the built-in scoring table, 28x28 int8 in ncbistdaa order
// (0 '-', 1 A, 2 B, 3 C, 4 D, 5 E, 6 F, 7 G, 8 H, 9 I, 10 K, 11 L, 12 M,
// 13 N, 14 P, 15 Q, 16 R, 17 S, 18 T, 19 V, 20 W, 21 X, 22 Y, 23 Z, 24 U,
// 25 '*', 26 O, 27 J), row = query residue.  The values are those of the
// table the reference stages on its GPU (gpu_blastp.c:46-75): the published
// BLOSUM62 for the 20 standard residues, NCBI's rows for B, Z and J, -1 for
// X, U and O, -4 against '*' ('*' vs '*' = 1) and -128 for the gap code.
#include <cstdint>
#include <cstring>
#include <mutex>

namespace bpx {

namespace {

// Lower triangle of the 20x20 core, rows/cols in ncbistdaa order of the
// standard residues: A C D E F G H I K L M N P Q R S T V W Y.
constexpr char kStd[] = "ACDEFGHIKLMNPQRSTVWY";
constexpr int8_t kTri[] = {
    4,                                                              // A
    0, 9,                                                           // C
    -2, -3, 6,                                                      // D
    -1, -4, 2, 5,                                                   // E
    -2, -2, -3, -3, 6,                                              // F
    0, -3, -1, -2, -3, 6,                                           // G
    -2, -3, -1, 0, -1, -2, 8,                                       // H
    -1, -1, -3, -3, 0, -4, -3, 4,                                   // I
    -1, -3, -1, 1, -3, -2, -1, -3, 5,                               // K
    -1, -1, -4, -3, 0, -4, -3, 2, -2, 4,                            // L
    -1, -1, -3, -2, 0, -3, -2, 1, -1, 2, 5,                         // M
    -2, -3, 1, 0, -3, 0, 1, -3, 0, -3, -2, 6,                       // N
    -1, -3, -1, -1, -4, -2, -2, -3, -1, -3, -2, -2, 7,              // P
    -1, -3, 0, 2, -3, -2, 0, -3, 1, -2, 0, 0, -1, 5,                // Q
    -1, -3, -2, 0, -3, -2, 0, -3, 2, -2, -1, 0, -2, 1, 5,           // R
    1, -1, 0, 0, -2, 0, -1, -2, 0, -2, -1, 1, -1, 0, -1, 4,         // S
    0, -1, -1, -1, -2, -2, -2, -1, -1, -1, -1, 0, -1, -1, -1, 1, 5, // T
    0, -1, -3, -2, -1, -3, -3, 3, -2, 1, 1, -3, -2, -2, -3, -2, 0, 4,          // V
    -3, -2, -4, -3, 1, -2, -2, -3, -3, -2, -1, -4, -4, -2, -3, -3, -2, -3, 11, // W
    -2, -2, -3, -2, 3, -3, 2, -1, -2, -1, -1, -2, -3, -1, -2, -2, -2, -1, 2, 7 // Y
};
// NCBI ambiguity rows, against "ACDEFGHIKLMNPQRSTVWY" then B Z J.
constexpr int8_t kB[23] = {-2, -3, 4, 1, -3, -1, 0, -3, 0, -4, -3, 4, -2, 0, -1, 0, -1, -3, -4, -3, 4, 0, -3};
constexpr int8_t kZ[23] = {-1, -3, 1, 4, -3, -2, 0, -3, 1, -3, -1, 0, -1, 4, 0, 0, -1, -2, -2, -2, 0, 4, -3};
constexpr int8_t kJ[23] = {-1, -1, -3, -3, 0, -4, -3, 3, -3, 3, 2, -3, -3, -2, -2, -2, -1, 2, -2, -1, -3, -3, 3};

constexpr char kStdaa[] = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";

int8_t g_table[28 * 28];
std::once_flag g_once;

int std_index(char c) {
    const char* p = std::strchr(kStd, c);
    return p ? (int)(p - kStd) : -1;
}

int8_t amb_score(const int8_t* row, char other) {
    const int i = std_index(other);
    if (i >= 0) return row[i];
    if (other == 'B') return row[20];
    if (other == 'Z') return row[21];
    return row[22];  // 'J'
}

void init() {
    for (int i = 0; i < 28; ++i)
        for (int j = 0; j < 28; ++j) {
            const char a = kStdaa[i], b = kStdaa[j];
            int8_t v;
            if (a == '-' || b == '-') {
                v = -128;
            } else if (a == '*' || b == '*') {
                v = (a == b) ? 1 : -4;
            } else if (std::strchr("XUO", a) || std::strchr("XUO", b)) {
                v = -1;
            } else if (a == 'B' || a == 'Z' || a == 'J') {
                v = amb_score(a == 'B' ? kB : a == 'Z' ? kZ : kJ, b);
            } else if (b == 'B' || b == 'Z' || b == 'J') {
                v = amb_score(b == 'B' ? kB : b == 'Z' ? kZ : kJ, a);
            } else {
                int x = std_index(a), y = std_index(b);
                if (x < y) { int t = x; x = y; y = t; }
                v = kTri[x * (x + 1) / 2 + y];
            }
            g_table[i * 28 + j] = v;
        }
}

}  // namespace

const int8_t* blosum62() {
    std::call_once(g_once, init);
    return g_table;
}

}  // namespace bpx
