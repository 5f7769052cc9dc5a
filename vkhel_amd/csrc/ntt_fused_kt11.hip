// Fused NTT kernels with 2048-word tiles for the two-pass sizes.
#include "ntt_fused.h"

namespace vk {

static const std::array<FusedEntry, kFusedEntries> g_table =
	make_fused_table<11>(std::make_integer_sequence<int, kFusedEntries>());

const FusedEntry *fused_table_kt11() {
	return g_table.data();
}

}  // namespace vk
