// pp_net.h -- host side of a proposal network (pp_net.cpp, pp_api.cpp).
#pragma once

#include <vector>

#include "pp_internal.h"

struct pp_net {
	int device = 0;
	hipStream_t stream = nullptr;
	ppd::NetDev d{};
	int step = 0;  // Adam steps taken
	// device scratch of host-driven calls and of training batches
	int cap = 0;  // rows (multiple of 32)
	__bf16 *x = nullptr, *h = nullptr, *dy = nullptr, *dh = nullptr;
	float *y = nullptr, *target = nullptr, *loss = nullptr;
	std::vector<void *> allocs;
	~pp_net();
};

namespace ppd {
// grow the scratch of `net` to at least m rows; PP_OK or an error status
int net_reserve(pp_net *net, int m);
}
