// mci_fast.hip — placeholder: the specialised kernels are not enabled yet.
#include "mci_fast.hpp"

namespace mcid
{
bool fast_eligible(int, int, int, int, int, int, int, int) { return false; }
int fast_sample(FastState &, const Chain &, int, int64_t, const ObsDev *, int, hipStream_t) { return MCI_E_RUNTIME; }
void fast_release(FastState &) {}
} // namespace mcid
