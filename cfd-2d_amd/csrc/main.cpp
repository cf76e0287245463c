// cfd2d_run: the driver of the reference binary (src/main.cpp:10-39) --
// opens ./task.log, builds the method named in ./task.xml, runs it.
#include <cstdio>

#include "method.hpp"

int main(int argc, char** argv)
{
	char defaultTask[] = "task.xml";
	char* task = argc > 1 ? argv[1] : defaultTask;
	cfd2d::hLog = fopen("task.log", "w");
	cfd2d::Method* m = cfd2d::Solver::initMethod(task);
	cfd2d::Solver::runMethod(m);
	cfd2d::Solver::destroyMethod(m);
	if (cfd2d::hLog) fclose(cfd2d::hLog);
	return 0;
}
