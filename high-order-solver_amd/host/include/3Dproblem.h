// 3Dproblem.h — reference header name; the 3D call surface is declared in gks_host.h
#pragma once
#include "gks_host.h"
