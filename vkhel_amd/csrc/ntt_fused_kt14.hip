// Fused NTT kernels with 16384-word tiles (144 KiB of LDS, one workgroup per CU) for the two-pass sizes.
#include "ntt_fused.h"

namespace vk {

static const std::array<FusedEntry, kFusedEntries> g_table =
	make_fused_table<14>(std::make_integer_sequence<int, kFusedEntries>());

const FusedEntry *fused_table_kt14() {
	return g_table.data();
}

}  // namespace vk
