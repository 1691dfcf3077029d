#pragma once
#include "../hipemu.h"
