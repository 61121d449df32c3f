// Size of the parameter-selection network's parameter vector.
#pragma once
#define MLP_NPARAM (4 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4)
