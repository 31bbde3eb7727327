// Version of the portableRT API this backend implements (version.hpp.in in the reference) and
// of the MI355X backend itself.
#pragma once
#define PORTABLERT_VERSION "0.3.3"
#define PORTABLERT_VERSION_MAJOR 0
#define PORTABLERT_VERSION_MINOR 3
#define PORTABLERT_VERSION_PATCH 3
#define PORTABLERT_MI355X_VERSION "0.1.0"
