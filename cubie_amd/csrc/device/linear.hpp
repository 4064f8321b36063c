// Placeholder: the linear component is added with the implicit families.
#pragma once
