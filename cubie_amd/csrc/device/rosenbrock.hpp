// Placeholder: the rosenbrock component is added with the implicit families.
#pragma once
