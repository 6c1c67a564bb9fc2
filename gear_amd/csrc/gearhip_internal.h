// Internal include: the public C ABI plus helpers shared by the .hip files.
#pragma once
#include "../../include/gearhip.h"
