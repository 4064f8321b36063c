// Placeholder: the firk component is added with the implicit families.
#pragma once
