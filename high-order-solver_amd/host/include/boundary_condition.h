// boundary_condition.h — reference header name; the 2D call surface is declared in gks_host.h
#pragma once
#include "gks_host.h"
