// sehf_psi.h -- per-energy files of the back-propagated wave function psi_n
// (host side of sehf_scan_set_psi_store / sehf_scan_dipole_only).
//
// One directory per energy, <root>/psi_ik<NNNN> (NNNN: the energy's index in
// the k list of the call), holding
//   psi_asym.bin  the asymptotic tail psi_n, n in [N - n_asym + 1, N]
//                 (n_asym = min(300, Nr)), in the reference's single-file
//                 layout (BackPropagator.cpp:27-101): a 32-byte header
//                 'PSI_ASYM', int32 version 1, int32 n_asym, int32 first
//                 index, int32 N_psi, 8 reserved bytes, then the matrices in
//                 ascending n, each N_psi x N_psi column-major;
//   psi_main.bin  psi_n for n in [0, n_main_hi] (the range the dipole
//                 integrals read, main.cpp:1259-1270) in the same layout with
//                 the magic 'PSI_MAIN' (the reference keeps this range in its
//                 chunked potential store instead);
//   __SUCCESS__   written last, after both files are complete and renamed
//                 into place: "k <value>" and the sizes, one per line.
// A file is written under a temporary name and renamed, so a run that fails
// leaves no marker and no file that looks complete.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sehf {

constexpr int kPsiTailPoints = 300;   // BackPropagator::Config::n_asym (main.cpp:1324)

struct PsiLayout {
    int Np = 0, Nr = 0;
    int tail_lo = 0, n_tail = 0;    // psi_asym.bin: [tail_lo, tail_lo + n_tail)
    int n_main = 0;                 // psi_main.bin: [0, n_main)
};

std::string psi_dir(const std::string& root, int ik);

// Writes one batch's files.  put() takes the batch's matrices of point n
// (nb consecutive N_psi x N_psi blocks); commit() completes every energy.
class PsiWriter {
   public:
    PsiWriter(const std::string& root, int ik0, int nb, const double* ks, const PsiLayout& lay, bool tail,
              bool main);
    ~PsiWriter();
    PsiWriter(const PsiWriter&) = delete;
    PsiWriter& operator=(const PsiWriter&) = delete;
    bool wants(int file, int n) const;   // file 0: tail, 1: main
    void put(int file, int n, const double* batch);
    void commit();

   private:
    void open_all(const std::string& root, int ik0, const double* ks);
    void cleanup();
    struct One {
        std::string dir;
        int fd[2] = {-1, -1};
        std::string tmp[2], fin[2];
        double k = 0.0;
    };
    std::vector<One> e_;
    PsiLayout lay_;
    bool on_[2] = {false, false};
    bool done_ = false;
};

// Reads one batch's files back for a replay of the fit and the dipoles.
class PsiReader {
   public:
    // need_main: psi_main.bin must hold [0, n_main) of this layout.
    PsiReader(const std::string& root, int ik0, int nb, const double* ks, const PsiLayout& lay, bool need_main);
    ~PsiReader();
    PsiReader(const PsiReader&) = delete;
    PsiReader& operator=(const PsiReader&) = delete;
    // psi_n of every energy of the batch into nb consecutive blocks: from the
    // main file where it holds n, else from the tail.
    void get(int n, double* batch) const;

   private:
    void open_all(const std::string& root, int ik0, const double* ks);
    void close_all();
    struct One {
        int fd[2] = {-1, -1};
        std::string path[2];
    };
    std::vector<One> e_;
    PsiLayout lay_;
    bool main_ = false;
};

}  // namespace sehf
