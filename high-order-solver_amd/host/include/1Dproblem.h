// 1Dproblem.h — reference header name; the 1D call surface is declared in gks_host.h
#pragma once
#include "gks_host.h"
